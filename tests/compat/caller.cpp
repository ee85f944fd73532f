// A C++ program written against the reference's host interface only: it
// includes kronmult_compat.hpp (the reference's kronmultN_batched /
// _pbatched / _xbatched names and argument lists, kronmult6_batched.cpp:4-5,
// kronmult6_pbatched.cpp:4-6, kronmult6_xbatched.cpp:4-7, the vbatched
// template's arguments kronmult6_vbatched.hpp:10-16, and the batched slender
// GEMMs kgemm_nn_batched.cpp:5-10 and copies) and the HIP runtime, allocates
// every buffer and pointer array on the device as the reference's tests do
// (test_kronmult6_xbatched.cpp:104-133), and checks each call against an
// explicit Kronecker matrix (or product) built on the host.
//
// Every overload is called: kronmult{1..6}_{batched,pbatched,xbatched,
// vbatched} and kgemm_{nn,nt,tt}_batched.  Exit status 0 when every output is
// within 1e-13 (normwise) of the host result; prints one line per call.
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

unsigned g_seed = 12345u;

double urand()
{
  g_seed = g_seed * 1664525u + 1013904223u;
  return (g_seed >> 8) * (2.0 / 16777216.0) - 1.0;
}

std::vector<double> rvec(size_t n)
{
  std::vector<double> v(n);
  for (double &x : v)
    x = urand();
  return v;
}

// y += kron(A_1, ..., A_d) x; A_k column-major m_k x n_k with leading
// dimension ld_k; vector indices (i_1, ..., i_d) with i_d fastest
void host_kron(int d, const int *m, const int *n, const double *const *A,
               const int *ld, const double *x, double *y)
{
  long rows = 1, cols = 1;
  for (int k = 0; k < d; k++)
  {
    rows *= m[k];
    cols *= n[k];
  }
  for (long row = 0; row < rows; row++)
  {
    double s = 0;
    for (long col = 0; col < cols; col++)
    {
      double kv = 1;
      long r = row, c = col;
      for (int k = d - 1; k >= 0; k--)
      {
        kv *= A[k][(r % m[k]) + (c % n[k]) * ld[k]];
        r /= m[k];
        c /= n[k];
      }
      s += kv * x[col];
    }
    y[row] += s;
  }
}

template<typename V>
V *to_device(const std::vector<V> &h)
{
  V *p = nullptr;
  CHECK(hipMalloc(&p, sizeof(V) * (h.empty() ? 1 : h.size())));
  if (!h.empty())
    CHECK(hipMemcpy(p, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice));
  return p;
}

std::vector<double> from_device(const double *p, size_t n)
{
  std::vector<double> h(n);
  CHECK(hipMemcpy(h.data(), p, sizeof(double) * n, hipMemcpyDeviceToHost));
  return h;
}

double rel_err(const std::vector<double> &got, const std::vector<double> &want)
{
  double e = 0, mx = 1;
  for (size_t i = 0; i < got.size(); i++)
  {
    e = std::fmax(e, std::fabs(got[i] - want[i]));
    mx = std::fmax(mx, std::fabs(want[i]));
  }
  return e / mx;
}

int g_failures = 0;

void report(const char *what, int d, double e, bool x_kept = true)
{
  std::printf("kronmult%d_%-9s rel err %.3e%s\n", d, what, e, x_kept ? "" : ", X MODIFIED");
  g_failures += !(e <= 1e-13) + !x_kept;
}

// the reference-named overloads, selected by d
void call_batched(int d, int n, const double *A, double *X, double *Y, double *W, int b)
{
  switch (d)
  {
  case 1: kronmult1_batched(n, A, X, Y, W, b); break;
  case 2: kronmult2_batched(n, A, X, Y, W, b); break;
  case 3: kronmult3_batched(n, A, X, Y, W, b); break;
  case 4: kronmult4_batched(n, A, X, Y, W, b); break;
  case 5: kronmult5_batched(n, A, X, Y, W, b); break;
  default: kronmult6_batched(n, A, X, Y, W, b); break;
  }
}
void call_pbatched(int d, int n, const double *A, double **pX, double **pY, double **pW, int b)
{
  switch (d)
  {
  case 1: kronmult1_pbatched(n, A, pX, pY, pW, b); break;
  case 2: kronmult2_pbatched(n, A, pX, pY, pW, b); break;
  case 3: kronmult3_pbatched(n, A, pX, pY, pW, b); break;
  case 4: kronmult4_pbatched(n, A, pX, pY, pW, b); break;
  case 5: kronmult5_pbatched(n, A, pX, pY, pW, b); break;
  default: kronmult6_pbatched(n, A, pX, pY, pW, b); break;
  }
}
void call_xbatched(int d, int n, const double *const *tab, int lda, double **pX, double **pY,
                   double **pW, int b)
{
  switch (d)
  {
  case 1: kronmult1_xbatched(n, tab, lda, pX, pY, pW, b); break;
  case 2: kronmult2_xbatched(n, tab, lda, pX, pY, pW, b); break;
  case 3: kronmult3_xbatched(n, tab, lda, pX, pY, pW, b); break;
  case 4: kronmult4_xbatched(n, tab, lda, pX, pY, pW, b); break;
  case 5: kronmult5_xbatched(n, tab, lda, pX, pY, pW, b); break;
  default: kronmult6_xbatched(n, tab, lda, pX, pY, pW, b); break;
  }
}
void call_vbatched(int d, const int *m, const int *n, const double *const *tab, double **pX,
                   double **pY, double *W, size_t cap, int b)
{
  switch (d)
  {
  case 1: kronmult1_vbatched(m[0], n[0], tab, pX, pY, W, cap, b); break;
  case 2: kronmult2_vbatched(m[0], n[0], m[1], n[1], tab, pX, pY, W, cap, b); break;
  case 3: kronmult3_vbatched(m[0], n[0], m[1], n[1], m[2], n[2], tab, pX, pY, W, cap, b); break;
  case 4: kronmult4_vbatched(m[0], n[0], m[1], n[1], m[2], n[2], m[3], n[3], tab, pX, pY, W, cap, b); break;
  case 5:
    kronmult5_vbatched(m[0], n[0], m[1], n[1], m[2], n[2], m[3], n[3], m[4], n[4], tab, pX, pY, W, cap, b);
    break;
  default:
    kronmult6_vbatched(m[0], n[0], m[1], n[1], m[2], n[2], m[3], n[3], m[4], n[4], m[5], n[5], tab, pX,
                       pY, W, cap, b);
    break;
  }
}

long ipow(int b, int e)
{
  long p = 1;
  for (int i = 0; i < e; i++)
    p *= b;
  return p;
}

void test_batched(int d, int n)
{
  const int B = 5;
  const long S = ipow(n, d);
  std::vector<double> A = rvec(n * n * d * B), X = rvec(S * B), Y = rvec(S * B);
  std::vector<double> want = Y;
  std::vector<int> mm(d, n), ld(d, n);
  for (int b = 0; b < B; b++)
  {
    std::vector<const double *> Ab(d);
    for (int k = 0; k < d; k++)
      Ab[k] = &A[(k + d * b) * n * n];
    host_kron(d, mm.data(), mm.data(), Ab.data(), ld.data(), &X[S * b], &want[S * b]);
  }
  double *dA = to_device(A), *dX = to_device(X), *dY = to_device(Y);
  double *dW = to_device(std::vector<double>(S * B, 0.0));
  call_batched(d, n, dA, dX, dY, dW, B);
  CHECK(hipDeviceSynchronize());
  report("batched", d, rel_err(from_device(dY, S * B), want), from_device(dX, S * B) == X);
  CHECK(hipFree(dA)); CHECK(hipFree(dX)); CHECK(hipFree(dY)); CHECK(hipFree(dW));
}

void test_pbatched(int d, int n)
{
  const int B = 7, NY = 3;
  const long S = ipow(n, d);
  std::vector<double> A = rvec(n * n * d * B), X = rvec(S * B), Y = rvec(S * NY);
  std::vector<double> want = Y;
  std::vector<int> mm(d, n), ld(d, n);
  for (int b = 0; b < B; b++)
  {
    std::vector<const double *> Ab(d);
    for (int k = 0; k < d; k++)
      Ab[k] = &A[(k + d * b) * n * n];
    host_kron(d, mm.data(), mm.data(), Ab.data(), ld.data(), &X[S * b], &want[S * (b % NY)]);
  }
  double *dA = to_device(A), *dX = to_device(X), *dY = to_device(Y);
  double *dW = to_device(std::vector<double>(S * B, 0.0));
  std::vector<double *> pX(B), pY(B), pW(B);
  for (int b = 0; b < B; b++)
  {
    pX[b] = dX + S * b;
    pY[b] = dY + S * (b % NY);
    pW[b] = dW + S * b;
  }
  double **dpX = to_device(pX), **dpY = to_device(pY), **dpW = to_device(pW);
  call_pbatched(d, n, dA, dpX, dpY, dpW, B);
  CHECK(hipDeviceSynchronize());
  report("pbatched", d, rel_err(from_device(dY, S * NY), want), from_device(dX, S * B) == X);
  CHECK(hipFree(dA)); CHECK(hipFree(dX)); CHECK(hipFree(dY)); CHECK(hipFree(dW));
  CHECK(hipFree(dpX)); CHECK(hipFree(dpY)); CHECK(hipFree(dpW));
}

// pooled A with lda = n + 2 through the (d, batch) pointer table, X shared
// between entries, default subbatchCount
void test_xbatched(int d, int n)
{
  const int B = 7, LDA = n + 2, POOL = 3, NX = 4, NY = 2;
  const long S = ipow(n, d);
  std::vector<double> pool = rvec(LDA * n * POOL), X = rvec(S * NX), Y = rvec(S * NY);
  double *dpool = to_device(pool), *dX = to_device(X), *dY = to_device(Y);
  double *dW = to_device(std::vector<double>(S * 2 * B, 0.0));
  std::vector<const double *> tab(d * B);
  std::vector<double *> pX(B), pY(B), pW(2 * B);
  std::vector<double> want = Y;
  std::vector<int> mm(d, n), ld(d, LDA);
  for (int b = 0; b < B; b++)
  {
    std::vector<const double *> Ab(d);
    for (int k = 0; k < d; k++)
    {
      const int p = (b + 2 * k) % POOL;
      tab[k + d * b] = dpool + p * LDA * n;
      Ab[k] = &pool[p * LDA * n];
    }
    host_kron(d, mm.data(), mm.data(), Ab.data(), ld.data(), &X[S * (b % NX)], &want[S * (b % NY)]);
    pX[b] = dX + S * (b % NX);
    pY[b] = dY + S * (b % NY);
  }
  for (int i = 0; i < 2 * B; i++)
    pW[i] = dW + S * i;
  const double **dtab = to_device(tab);
  double **dpX = to_device(pX), **dpY = to_device(pY), **dpW = to_device(pW);
  call_xbatched(d, n, dtab, LDA, dpX, dpY, dpW, B);
  CHECK(hipDeviceSynchronize());
  report("xbatched", d, rel_err(from_device(dY, S * NY), want), from_device(dX, S * NX) == X);
  CHECK(hipFree(dpool)); CHECK(hipFree(dX)); CHECK(hipFree(dY)); CHECK(hipFree(dW));
  CHECK(hipFree(dtab)); CHECK(hipFree(dpX)); CHECK(hipFree(dpY)); CHECK(hipFree(dpW));
}

// rectangular m_k x n_k factors (leading dimension m_k), outputs shared
void test_vbatched(int d, int n)
{
  const int B = 4, NY = 2;
  std::vector<int> m(d), nn(d);
  for (int k = 0; k < d; k++)
  {
    nn[k] = n + (k % 2);
    m[k] = n + ((k + 1) % 3) - 1;
  }
  long SX = 1, SY = 1, SW = 0;
  for (int k = 0; k < d; k++)
  {
    SX *= nn[k];
    SY *= m[k];
  }
  // the largest intermediate of either mode order (two slots per entry)
  for (int L = 1; L < d; L++)
  {
    long f = 1, bk = 1;
    for (int k = 0; k < d; k++)
    {
      f *= (k < L) ? m[k] : nn[k];
      bk *= (k >= d - L) ? m[k] : nn[k];
    }
    SW = std::max(SW, std::max(f, bk));
  }
  SW = std::max(SW, std::max(SX, SY));
  std::vector<std::vector<double>> mats(d * B);
  std::vector<double> X = rvec(SX * B), Y = rvec(SY * NY);
  std::vector<double> want = Y;
  std::vector<double *> dmats(d * B);
  for (int b = 0; b < B; b++)
  {
    std::vector<const double *> Ab(d);
    for (int k = 0; k < d; k++)
    {
      mats[k + d * b] = rvec(m[k] * nn[k]);
      dmats[k + d * b] = to_device(mats[k + d * b]);
      Ab[k] = mats[k + d * b].data();
    }
    host_kron(d, m.data(), nn.data(), Ab.data(), m.data(), &X[SX * b], &want[SY * (b % NY)]);
  }
  std::vector<const double *> tab(dmats.begin(), dmats.end());
  double *dX = to_device(X), *dY = to_device(Y);
  const size_t cap = sizeof(double) * 2 * SW * B;
  double *dW = to_device(std::vector<double>(2 * SW * B, 0.0));
  std::vector<double *> pX(B), pY(B);
  for (int b = 0; b < B; b++)
  {
    pX[b] = dX + SX * b;
    pY[b] = dY + SY * (b % NY);
  }
  const double **dtab = to_device(tab);
  double **dpX = to_device(pX), **dpY = to_device(pY);
  call_vbatched(d, m.data(), nn.data(), dtab, dpX, dpY, dW, cap, B);
  CHECK(hipDeviceSynchronize());
  report("vbatched", d, rel_err(from_device(dY, SY * NY), want), from_device(dX, SX * B) == X);
  for (double *p : dmats)
    CHECK(hipFree(p));
  CHECK(hipFree(dX)); CHECK(hipFree(dY)); CHECK(hipFree(dW));
  CHECK(hipFree(dtab)); CHECK(hipFree(dpX)); CHECK(hipFree(dpY));
}

// C_b = alpha op(A_b) op(B_b) + beta C_b; beta = 1 with a C shared by every
// entry (atomic adds), beta = 0 and beta = 0.5 with a C per entry
void test_kgemm(int op)
{
  const int MM = 37, NN = 5, KK = 4, B = 3;
  const double alpha = 1.5;
  const char *name[] = {"nn", "nt", "tt"};
  for (double beta : {0.0, 1.0, 0.5})
  {
    const bool shared = beta == 1.0;
    const int ra = op == 2 ? KK : MM, ca = op == 2 ? MM : KK; // A rows, cols
    const int rb = op == 0 ? KK : NN, cb = op == 0 ? NN : KK; // B rows, cols
    const int lda = ra + 1, ldb = rb + 2, ldc = MM + 3;
    const int nc = shared ? 1 : B;
    std::vector<double> A = rvec(lda * ca * B), Bm = rvec(ldb * cb * B), C = rvec(ldc * NN * nc);
    std::vector<double> want = C;
    for (int c = 0; c < nc; c++)
      for (int j = 0; j < NN; j++)
        for (int i = 0; i < MM; i++)
          want[c * ldc * NN + i + j * ldc] *= beta;
    for (int b = 0; b < B; b++)
    {
      const int c = shared ? 0 : b;
      for (int j = 0; j < NN; j++)
        for (int i = 0; i < MM; i++)
        {
          double s = 0;
          for (int k = 0; k < KK; k++)
          {
            const double a = op == 2 ? A[b * lda * ca + k + i * lda] : A[b * lda * ca + i + k * lda];
            const double bv = op == 0 ? Bm[b * ldb * cb + k + j * ldb] : Bm[b * ldb * cb + j + k * ldb];
            s += a * bv;
          }
          want[c * ldc * NN + i + j * ldc] += alpha * s;
        }
    }
    double *dA = to_device(A), *dB = to_device(Bm), *dC = to_device(C);
    std::vector<double *> pA(B), pB(B), pC(B);
    std::vector<int> lA(B, lda), lB(B, ldb), lC(B, ldc);
    for (int b = 0; b < B; b++)
    {
      pA[b] = dA + b * lda * ca;
      pB[b] = dB + b * ldb * cb;
      pC[b] = dC + (shared ? 0 : b * ldc * NN);
    }
    double **dpA = to_device(pA), **dpB = to_device(pB), **dpC = to_device(pC);
    int *dlA = to_device(lA), *dlB = to_device(lB), *dlC = to_device(lC);
    if (op == 0)
      kgemm_nn_batched(MM, NN, KK, alpha, dpA, dlA, dpB, dlB, beta, dpC, dlC, B);
    else if (op == 1)
      kgemm_nt_batched(MM, NN, KK, alpha, dpA, dlA, dpB, dlB, beta, dpC, dlC, B);
    else
      kgemm_tt_batched(MM, NN, KK, alpha, dpA, dlA, dpB, dlB, beta, dpC, dlC, B);
    CHECK(hipDeviceSynchronize());
    const double e = rel_err(from_device(dC, ldc * NN * nc), want);
    std::printf("kgemm_%s_batched beta %.1f rel err %.3e\n", name[op], beta, e);
    g_failures += !(e <= 1e-13);
    CHECK(hipFree(dA)); CHECK(hipFree(dB)); CHECK(hipFree(dC));
    CHECK(hipFree(dpA)); CHECK(hipFree(dpB)); CHECK(hipFree(dpC));
    CHECK(hipFree(dlA)); CHECK(hipFree(dlB)); CHECK(hipFree(dlC));
  }
}

} // namespace

int main()
{
  // n per d keeps the host's explicit Kronecker matrix small
  const int nd[7] = {0, 7, 5, 4, 4, 3, 3};
  for (int d = 1; d <= 6; d++)
  {
    test_batched(d, nd[d]);
    test_pbatched(d, nd[d]);
    test_xbatched(d, nd[d]);
    test_vbatched(d, nd[d]);
  }
  for (int op = 0; op < 3; op++)
    test_kgemm(op);
  std::printf("%s\n", g_failures ? "FAILED" : "ok");
  return g_failures ? 1 : 0;
}
