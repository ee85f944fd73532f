// kronmult_compat.hpp -- drop-in C++ overloads with the reference's host
// wrapper names and argument lists, forwarding to the C ABI (kronmult_amd.h)
// on the default stream.  Like the reference wrappers they return void and
// launch asynchronously; unlike them they check their arguments: a call the
// library rejects prints the status on stderr and aborts (the reference
// would run on and corrupt memory or silently compute nothing).
//
// Replaced reference functions:
//   kronmultN_batched   kronmult6_batched.cpp:4-18    (N = 1..6)
//   kronmultN_pbatched  kronmult6_pbatched.cpp:4-19
//   kronmultN_xbatched  kronmult6_xbatched.cpp:4-28   (default subbatchCount = 0)
//   kronmultN_vbatched  kronmult6_vbatched.hpp:10-34  (a __global__ template in
//                       the reference; here a host function with its arguments)
// Size limit: shapes served by the workspace-staged kernel (rectangular
// vbatched factors, and square shapes without an on-chip kernel) take
// n_k <= 300; a larger n_k is reported as "not supported" and the call
// aborts.  The reference has no such limit.
#ifndef KRONMULT_COMPAT_HPP
#define KRONMULT_COMPAT_HPP

#include <cstddef>

void kronmult1_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount);
void kronmult2_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount);
void kronmult3_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount);
void kronmult4_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount);
void kronmult5_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount);
void kronmult6_batched(int const n, double const Aarray_[], double Xarray_[],
                       double Yarray_[], double Warray_[], int const batchCount);
void kronmult1_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount);
void kronmult2_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount);
void kronmult3_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount);
void kronmult4_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount);
void kronmult5_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount);
void kronmult6_pbatched(int const n, double const Aarray_[], double *Xarray_[],
                        double *Yarray_[], double *Warray_[],
                        int const batchCount);
void kronmult1_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount = 0);
void kronmult2_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount = 0);
void kronmult3_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount = 0);
void kronmult4_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount = 0);
void kronmult5_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount = 0);
void kronmult6_xbatched(int const n, double const *const Aarray_[],
                        int const lda, double *Xarray_[], double *Yarray_[],
                        double *Warray_[], int const batchCount,
                        int const subbatchCount = 0);
void kronmult1_vbatched(int const m1, int const n1,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount);
void kronmult2_vbatched(int const m1, int const n1, int const m2, int const n2,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount);
void kronmult3_vbatched(int const m1, int const n1, int const m2, int const n2, int const m3, int const n3,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount);
void kronmult4_vbatched(int const m1, int const n1, int const m2, int const n2, int const m3, int const n3, int const m4, int const n4,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount);
void kronmult5_vbatched(int const m1, int const n1, int const m2, int const n2, int const m3, int const n3, int const m4, int const n4, int const m5, int const n5,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount);
void kronmult6_vbatched(int const m1, int const n1, int const m2, int const n2, int const m3, int const n3, int const m4, int const n4, int const m5, int const n5, int const m6, int const n6,
                        double const *const Aarray_[], double *pX_[],
                        double *pY_[], double *W_, size_t const Wcapacity,
                        int const batchCount);

// batched slender GEMMs (kgemm_nn_batched.cpp:5-10, kgemm_nt_batched.cpp:5-10,
// kgemm_tt_batched.cpp:5-10)
void kgemm_nn_batched(int const mm, int const nn, int const kk,
                      double const alpha, double *const Aarray_[],
                      int const ldAarray_[], double *const Barray_[],
                      int const ldBarray_[], double const beta,
                      double *const Carray_[], int const ldCarray_[],
                      int const batchCount);
void kgemm_nt_batched(int const mm, int const nn, int const kk,
                      double const alpha, double *const Aarray_[],
                      int const ldAarray_[], double *const Barray_[],
                      int const ldBarray_[], double const beta,
                      double *const Carray_[], int const ldCarray_[],
                      int const batchCount);
void kgemm_tt_batched(int const mm, int const nn, int const kk,
                      double const alpha, double *const Aarray_[],
                      int const ldAarray_[], double *const Barray_[],
                      int const ldBarray_[], double const beta,
                      double *const Carray_[], int const ldCarray_[],
                      int const batchCount);

#endif
