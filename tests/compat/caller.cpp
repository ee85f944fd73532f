// A C++ program written against the reference's host interface only: it
// includes kronmult_compat.hpp (the reference's kronmultN_batched /
// _pbatched / _xbatched names and argument lists, kronmult6_batched.cpp:4-5,
// kronmult6_pbatched.cpp:4-6, kronmult6_xbatched.cpp:4-7) and the HIP
// runtime, allocates every buffer and pointer array on the device as the
// reference's tests do (test_kronmult6_xbatched.cpp:104-133), and checks each
// call against an explicit Kronecker matrix built on the host.
//
// Exit status 0 when every output is within 1e-13 (normwise) of the host
// result; prints one line per entry point.
#include "kronmult_compat.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(call)                                                            \
  do                                                                           \
  {                                                                            \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
    {                                                                          \
      std::fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));          \
      std::exit(2);                                                            \
    }                                                                          \
  } while (0)

namespace
{

const int D = 3, N = 4, S = 64; // S = N^D
const int BATCH = 7;

double urand(unsigned &s)
{
  s = s * 1664525u + 1013904223u;
  return (s >> 8) * (2.0 / 16777216.0) - 1.0;
}

// y += kron(A_1, ..., A_D) x; A_k column-major N x N with leading dimension
// lda; the vector index is (i_1, ..., i_D) with i_D fastest
void host_entry(const double *const A[D], int lda, const double *x, double *y)
{
  for (int row = 0; row < S; row++)
  {
    double s = 0;
    for (int col = 0; col < S; col++)
    {
      double k = 1;
      int r = row, c = col;
      for (int m = D - 1; m >= 0; m--)
      {
        k *= A[m][(r % N) + (c % N) * lda];
        r /= N;
        c /= N;
      }
      s += k * x[col];
    }
    y[row] += s;
  }
}

template<typename V>
V *to_device(const std::vector<V> &h)
{
  V *d = nullptr;
  CHECK(hipMalloc(&d, sizeof(V) * h.size()));
  CHECK(hipMemcpy(d, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice));
  return d;
}

double rel_err(const std::vector<double> &got, const std::vector<double> &want)
{
  double e = 0, m = 1;
  for (size_t i = 0; i < got.size(); i++)
  {
    e = std::fmax(e, std::fabs(got[i] - want[i]));
    m = std::fmax(m, std::fabs(want[i]));
  }
  return e / m;
}

} // namespace

int main()
{
  unsigned seed = 12345u;
  int failures = 0;

  // kronmult3_batched: A is (N, N, D, batch), X/Y/W columns of S
  {
    std::vector<double> A(N * N * D * BATCH), X(S * BATCH), Y(S * BATCH);
    for (double &v : A) v = urand(seed);
    for (double &v : X) v = urand(seed);
    for (double &v : Y) v = urand(seed);
    std::vector<double> want = Y;
    for (int b = 0; b < BATCH; b++)
    {
      const double *Ab[D];
      for (int k = 0; k < D; k++)
        Ab[k] = &A[(k + D * b) * N * N];
      host_entry(Ab, N, &X[S * b], &want[S * b]);
    }
    double *dA = to_device(A), *dX = to_device(X), *dY = to_device(Y);
    double *dW = to_device(std::vector<double>(S * BATCH, 0.0));
    kronmult3_batched(N, dA, dX, dY, dW, BATCH);
    CHECK(hipDeviceSynchronize());
    std::vector<double> got(S * BATCH);
    CHECK(hipMemcpy(got.data(), dY, sizeof(double) * got.size(), hipMemcpyDeviceToHost));
    double e = rel_err(got, want);
    std::printf("kronmult3_batched  rel err %.3e\n", e);
    failures += !(e <= 1e-13);
    CHECK(hipFree(dA)); CHECK(hipFree(dX)); CHECK(hipFree(dY)); CHECK(hipFree(dW));
  }

  // kronmult3_pbatched: device pointer arrays, outputs shared 3 ways
  {
    const int NY = 3;
    std::vector<double> A(N * N * D * BATCH), X(S * BATCH), Y(S * NY);
    for (double &v : A) v = urand(seed);
    for (double &v : X) v = urand(seed);
    for (double &v : Y) v = urand(seed);
    std::vector<double> want = Y;
    for (int b = 0; b < BATCH; b++)
    {
      const double *Ab[D];
      for (int k = 0; k < D; k++)
        Ab[k] = &A[(k + D * b) * N * N];
      host_entry(Ab, N, &X[S * b], &want[S * (b % NY)]);
    }
    double *dA = to_device(A), *dX = to_device(X), *dY = to_device(Y);
    double *dW = to_device(std::vector<double>(S * BATCH, 0.0));
    std::vector<double *> pX(BATCH), pY(BATCH), pW(BATCH);
    for (int b = 0; b < BATCH; b++)
    {
      pX[b] = dX + S * b;
      pY[b] = dY + S * (b % NY);
      pW[b] = dW + S * b;
    }
    double **dpX = to_device(pX), **dpY = to_device(pY), **dpW = to_device(pW);
    kronmult3_pbatched(N, dA, dpX, dpY, dpW, BATCH);
    CHECK(hipDeviceSynchronize());
    std::vector<double> got(S * NY);
    CHECK(hipMemcpy(got.data(), dY, sizeof(double) * got.size(), hipMemcpyDeviceToHost));
    double e = rel_err(got, want);
    std::printf("kronmult3_pbatched rel err %.3e\n", e);
    failures += !(e <= 1e-13);
    CHECK(hipFree(dA)); CHECK(hipFree(dX)); CHECK(hipFree(dY)); CHECK(hipFree(dW));
    CHECK(hipFree(dpX)); CHECK(hipFree(dpY)); CHECK(hipFree(dpW));
  }

  // kronmult3_xbatched: pooled A with lda = N + 2 through a (D, batch)
  // pointer table, X shared between entries, default subbatchCount
  {
    const int LDA = N + 2, POOL = 3, NX = 4, NY = 2;
    std::vector<double> pool(LDA * N * POOL), X(S * NX), Y(S * NY);
    for (double &v : pool) v = urand(seed);
    for (double &v : X) v = urand(seed);
    for (double &v : Y) v = urand(seed);
    double *dpool = to_device(pool), *dX = to_device(X), *dY = to_device(Y);
    double *dW = to_device(std::vector<double>(S * 2 * BATCH, 0.0));
    std::vector<const double *> tab(D * BATCH);
    std::vector<double *> pX(BATCH), pY(BATCH), pW(2 * BATCH);
    std::vector<double> want = Y;
    for (int b = 0; b < BATCH; b++)
    {
      const double *Ab[D];
      for (int k = 0; k < D; k++)
      {
        const int p = (b + 2 * k) % POOL;
        tab[k + D * b] = dpool + p * LDA * N;
        Ab[k] = &pool[p * LDA * N];
      }
      host_entry(Ab, LDA, &X[S * (b % NX)], &want[S * (b % NY)]);
      pX[b] = dX + S * (b % NX);
      pY[b] = dY + S * (b % NY);
    }
    for (int i = 0; i < 2 * BATCH; i++)
      pW[i] = dW + S * i;
    const double **dtab = to_device(tab);
    double **dpX = to_device(pX), **dpY = to_device(pY), **dpW = to_device(pW);
    kronmult3_xbatched(N, dtab, LDA, dpX, dpY, dpW, BATCH);
    CHECK(hipDeviceSynchronize());
    std::vector<double> got(S * NY), xs(S * NX);
    CHECK(hipMemcpy(got.data(), dY, sizeof(double) * got.size(), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(xs.data(), dX, sizeof(double) * xs.size(), hipMemcpyDeviceToHost));
    double e = rel_err(got, want);
    const bool x_kept = xs == X;
    std::printf("kronmult3_xbatched rel err %.3e, X %s\n", e, x_kept ? "unchanged" : "MODIFIED");
    failures += !(e <= 1e-13) + !x_kept;
    CHECK(hipFree(dpool)); CHECK(hipFree(dX)); CHECK(hipFree(dY)); CHECK(hipFree(dW));
    CHECK(hipFree(dtab)); CHECK(hipFree(dpX)); CHECK(hipFree(dpY)); CHECK(hipFree(dpW));
  }

  std::printf("%s\n", failures ? "FAILED" : "ok");
  return failures ? 1 : 0;
}
