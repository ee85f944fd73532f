// km_compat.cpp -- the reference-named C++ overloads declared in
// include/kronmult_compat.hpp, each a thin forward to the C ABI.
#include <cstdio>
#include <cstdlib>

#include "../../include/kronmult_amd.h"
#include "../../include/kronmult_compat.hpp"

namespace
{
void check(km_status s, const char *fn)
{
  if (s != KM_SUCCESS)
  {
    std::fprintf(stderr, "kronmult_amd: %s failed: %s\n", fn,
                 km_status_string(s));
    std::abort();
  }
}
} // namespace

void kronmult1_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount)
{
  check(km_kronmult_batched_f64(1, n, Aarray_, Xarray_, Yarray_, Warray_,
                                batchCount, nullptr, nullptr),
        "kronmult1_batched");
}

void kronmult2_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount)
{
  check(km_kronmult_batched_f64(2, n, Aarray_, Xarray_, Yarray_, Warray_,
                                batchCount, nullptr, nullptr),
        "kronmult2_batched");
}

void kronmult3_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount)
{
  check(km_kronmult_batched_f64(3, n, Aarray_, Xarray_, Yarray_, Warray_,
                                batchCount, nullptr, nullptr),
        "kronmult3_batched");
}

void kronmult4_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount)
{
  check(km_kronmult_batched_f64(4, n, Aarray_, Xarray_, Yarray_, Warray_,
                                batchCount, nullptr, nullptr),
        "kronmult4_batched");
}

void kronmult5_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount)
{
  check(km_kronmult_batched_f64(5, n, Aarray_, Xarray_, Yarray_, Warray_,
                                batchCount, nullptr, nullptr),
        "kronmult5_batched");
}

void kronmult6_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount)
{
  check(km_kronmult_batched_f64(6, n, Aarray_, Xarray_, Yarray_, Warray_,
                                batchCount, nullptr, nullptr),
        "kronmult6_batched");
}

void kronmult1_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount)
{
  check(km_kronmult_pbatched_f64(1, n, Aarray_, Xarray_, Yarray_, Warray_,
                                 batchCount, nullptr, nullptr),
        "kronmult1_pbatched");
}

void kronmult2_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount)
{
  check(km_kronmult_pbatched_f64(2, n, Aarray_, Xarray_, Yarray_, Warray_,
                                 batchCount, nullptr, nullptr),
        "kronmult2_pbatched");
}

void kronmult3_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount)
{
  check(km_kronmult_pbatched_f64(3, n, Aarray_, Xarray_, Yarray_, Warray_,
                                 batchCount, nullptr, nullptr),
        "kronmult3_pbatched");
}

void kronmult4_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount)
{
  check(km_kronmult_pbatched_f64(4, n, Aarray_, Xarray_, Yarray_, Warray_,
                                 batchCount, nullptr, nullptr),
        "kronmult4_pbatched");
}

void kronmult5_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount)
{
  check(km_kronmult_pbatched_f64(5, n, Aarray_, Xarray_, Yarray_, Warray_,
                                 batchCount, nullptr, nullptr),
        "kronmult5_pbatched");
}

void kronmult6_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount)
{
  check(km_kronmult_pbatched_f64(6, n, Aarray_, Xarray_, Yarray_, Warray_,
                                 batchCount, nullptr, nullptr),
        "kronmult6_pbatched");
}

void kronmult1_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount)
{
  check(km_kronmult_xbatched_f64(1, n, Aarray_, lda, Xarray_, Yarray_,
                                 Warray_, batchCount, subbatchCount, nullptr,
                                 nullptr),
        "kronmult1_xbatched");
}

void kronmult2_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount)
{
  check(km_kronmult_xbatched_f64(2, n, Aarray_, lda, Xarray_, Yarray_,
                                 Warray_, batchCount, subbatchCount, nullptr,
                                 nullptr),
        "kronmult2_xbatched");
}

void kronmult3_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount)
{
  check(km_kronmult_xbatched_f64(3, n, Aarray_, lda, Xarray_, Yarray_,
                                 Warray_, batchCount, subbatchCount, nullptr,
                                 nullptr),
        "kronmult3_xbatched");
}

void kronmult4_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount)
{
  check(km_kronmult_xbatched_f64(4, n, Aarray_, lda, Xarray_, Yarray_,
                                 Warray_, batchCount, subbatchCount, nullptr,
                                 nullptr),
        "kronmult4_xbatched");
}

void kronmult5_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount)
{
  check(km_kronmult_xbatched_f64(5, n, Aarray_, lda, Xarray_, Yarray_,
                                 Warray_, batchCount, subbatchCount, nullptr,
                                 nullptr),
        "kronmult5_xbatched");
}

void kronmult6_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount)
{
  check(km_kronmult_xbatched_f64(6, n, Aarray_, lda, Xarray_, Yarray_,
                                 Warray_, batchCount, subbatchCount, nullptr,
                                 nullptr),
        "kronmult6_xbatched");
}

void kronmult1_vbatched(int const m1, int const n1,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount)
{
  int const m[1] = {m1};
  int const n[1] = {n1};
  check(km_kronmult_vbatched_f64(1, m, n, Aarray_, pX_, pY_, W_, Wcapacity,
                                 batchCount, nullptr, nullptr),
        "kronmult1_vbatched");
}

void kronmult2_vbatched(int const m1, int const n1, int const m2, int const n2,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount)
{
  int const m[2] = {m1, m2};
  int const n[2] = {n1, n2};
  check(km_kronmult_vbatched_f64(2, m, n, Aarray_, pX_, pY_, W_, Wcapacity,
                                 batchCount, nullptr, nullptr),
        "kronmult2_vbatched");
}

void kronmult3_vbatched(int const m1, int const n1, int const m2, int const n2, int const m3, int const n3,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount)
{
  int const m[3] = {m1, m2, m3};
  int const n[3] = {n1, n2, n3};
  check(km_kronmult_vbatched_f64(3, m, n, Aarray_, pX_, pY_, W_, Wcapacity,
                                 batchCount, nullptr, nullptr),
        "kronmult3_vbatched");
}

void kronmult4_vbatched(int const m1, int const n1, int const m2, int const n2, int const m3, int const n3, int const m4, int const n4,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount)
{
  int const m[4] = {m1, m2, m3, m4};
  int const n[4] = {n1, n2, n3, n4};
  check(km_kronmult_vbatched_f64(4, m, n, Aarray_, pX_, pY_, W_, Wcapacity,
                                 batchCount, nullptr, nullptr),
        "kronmult4_vbatched");
}

void kronmult5_vbatched(int const m1, int const n1, int const m2, int const n2, int const m3, int const n3, int const m4, int const n4, int const m5, int const n5,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount)
{
  int const m[5] = {m1, m2, m3, m4, m5};
  int const n[5] = {n1, n2, n3, n4, n5};
  check(km_kronmult_vbatched_f64(5, m, n, Aarray_, pX_, pY_, W_, Wcapacity,
                                 batchCount, nullptr, nullptr),
        "kronmult5_vbatched");
}

void kronmult6_vbatched(int const m1, int const n1, int const m2, int const n2, int const m3, int const n3, int const m4, int const n4, int const m5, int const n5, int const m6, int const n6,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount)
{
  int const m[6] = {m1, m2, m3, m4, m5, m6};
  int const n[6] = {n1, n2, n3, n4, n5, n6};
  check(km_kronmult_vbatched_f64(6, m, n, Aarray_, pX_, pY_, W_, Wcapacity,
                                 batchCount, nullptr, nullptr),
        "kronmult6_vbatched");
}

void kgemm_nn_batched(int const mm, int const nn, int const kk,
                      double const alpha, double *const Aarray_[],
                      int const ldAarray_[], double *const Barray_[],
                      int const ldBarray_[], double const beta,
                      double *const Carray_[], int const ldCarray_[],
                      int const batchCount)
{
  check(km_kgemm_nn_batched_f64(mm, nn, kk, alpha, Aarray_, ldAarray_, Barray_,
                                ldBarray_, beta, Carray_, ldCarray_, batchCount,
                                nullptr, nullptr),
        "kgemm_nn_batched");
}

void kgemm_nt_batched(int const mm, int const nn, int const kk,
                      double const alpha, double *const Aarray_[],
                      int const ldAarray_[], double *const Barray_[],
                      int const ldBarray_[], double const beta,
                      double *const Carray_[], int const ldCarray_[],
                      int const batchCount)
{
  check(km_kgemm_nt_batched_f64(mm, nn, kk, alpha, Aarray_, ldAarray_, Barray_,
                                ldBarray_, beta, Carray_, ldCarray_, batchCount,
                                nullptr, nullptr),
        "kgemm_nt_batched");
}

void kgemm_tt_batched(int const mm, int const nn, int const kk,
                      double const alpha, double *const Aarray_[],
                      int const ldAarray_[], double *const Barray_[],
                      int const ldBarray_[], double const beta,
                      double *const Carray_[], int const ldCarray_[],
                      int const batchCount)
{
  check(km_kgemm_tt_batched_f64(mm, nn, kk, alpha, Aarray_, ldAarray_, Barray_,
                                ldBarray_, beta, Carray_, ldCarray_, batchCount,
                                nullptr, nullptr),
        "kgemm_tt_batched");
}
