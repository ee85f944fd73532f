// km_internal.h -- host-side kernel registry shared by the instantiation
// units (km_lds_d*.hip, km_slab.hip, km_global.hip) and the C ABI (km_api.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <stddef.h>

#include "km_device.h"
#include "../../include/kronmult_amd.h"

// size of the allocation-free entry point's static per-device buffer
// (km_noalloc.hip): the device-built order of KM_NOALLOC_MAX_BATCH entries
// cut into up to kOrderMaxGroups ranges needs 7.1 MiB
#define KM_NOALLOC_SCRATCH_BYTES (8u << 20)

namespace km
{

enum Family
{
  FAM_LDS = 0,
  FAM_SLAB = 1,
  FAM_GLOBAL = 2,
  FAM_TWOPASS = 3,
  FAM_REG = 4,
  FAM_GROUPED = 5,
  FAM_HALF = 6,
  FAM_XCD = 7
};

struct KernelInfo
{
  const void *fn = nullptr; // __global__ function
  int nt = 0;               // threads per workgroup
  int family = FAM_LDS;
  int per_cu = 1;           // workgroups per CU of a cooperative launch
  int max_per_cu = 0;       // resident workgroups per CU the grid is sized
                            // for (0: as many as fit)
};

// threads per workgroup of kron_lds for a tensor of s elements
constexpr int lds_threads(int s) { return s <= 1024 ? 64 : (s <= 4096 ? 256 : 512); }
// wider workgroups for size bins that fit a few resident waves of
// workgroups, where one entry's latency bounds the launch (grouped calls)
constexpr int lds_threads_wide(int s) { return s <= 256 ? 64 : (s <= 2048 ? 256 : 512); }
// largest tensor kron_lds keeps in LDS (doubles): 150 KiB of the 160 KiB
constexpr int lds_max_elems(int elem_bytes) { return 150 * 1024 / elem_bytes; }

// kron_lds instantiation for (d, n) if one was built (n <= 8, n^d fits LDS)
template<typename T>
bool lds_kernel(int d, int n, bool exact, bool fwd, KernelInfo *out);
// per-d units
template<typename T, int D>
bool lds_kernel_d(int n, bool exact, bool fwd, KernelInfo *out);
// FMA kron_lds with lds_threads_wide threads, for the (d, n) where that
// differs from lds_threads and no kron_reg exists (km_lds_wide.hip)
template<typename T>
bool lds_kernel_wide(int d, int n, KernelInfo *out);

// batched slender GEMM (km_kgemm.hip); op 0 = nn, 1 = nt, 2 = tt
template<typename T>
struct KgemmArgs
{
  const T *const *A;
  const T *const *B;
  T *const *C;
  const int *lda;
  const int *ldb;
  const int *ldc;
  int mm, nn, kk, batch;
  int tiles; // row tiles per entry (set by the launcher)
  typename Acc<T>::type alpha, beta;
};
template<typename T>
hipError_t kgemm_launch(int op, KgemmArgs<T> &g, bool exact, int cus,
                        hipStream_t stream);

// pooled xbatched tables from index lists (km_tables.hip)
template<typename T>
struct TableArgs
{
  int d;
  long long batch;
  const T *pool[6];
  long long a_stride[6];
  int a_count[6];
  const int *aidx;
  const int *xidx;
  const int *yidx;
  T *X;
  long long x_stride;
  int x_count;
  T *Y;
  long long y_stride;
  int y_count;
  const T **Aarray;
  T **pX;
  T **pY;
  int *bad;
};
template<typename T>
hipError_t tables_launch(TableArgs<T> &t, int cus, hipStream_t stream);

#ifndef KM_USE_REG
#define KM_USE_REG 1
#endif
// kron_reg instantiation for (d, n) if one was built (forward mode order);
// all: also the ones that are not a default (tuned variants, owned bins)
template<typename T>
bool reg_kernel(int d, int n, bool exact, KernelInfo *out, bool all = false);

// kron_slab instantiation for (d, n) if one was built
template<typename T>
bool slab_kernel(int d, int n, bool exact, KernelInfo *out);

// Arguments of kron_runs and kron_runs_combine beyond EntryArgs (all device
// memory, from the group plan): nvec[b] = n of entry b; work unit u covers
// order positions [ustart[u], ustart[u+1]), a whole output run or one part
// of a split run; umeta[3u..3u+2] = split id (-1: whole run), part, parts;
// split run s keeps its part sums at ws + wsoff[s] + part*n^d, and its
// entries start at order position sfirst[s] (sparts[s] parts).
template<typename T>
struct RunsArgs
{
  const int *nvec;
  const int *ustart;
  const int *umeta;
  T *ws;
  const long long *wsoff;
  const int *sfirst;
  const int *sparts;
};

// kron_runs_combine: adds the part sums of every split run to its output
template<typename T>
const void *runs_combine_kernel(int d, int *nt);

// kron_runs (grouped variable-size batch in one launch, whole output runs
// per workgroup) for dimension d: largest n in *nmax, largest entry size
// (n^d) contracted by one wavefront in *small_max
template<typename T>
bool runs_kernel(int d, bool exact, KernelInfo *out, int *nmax, int *small_max);

// kron_half (two workgroups per entry, split by the last mode's rows; FMA
// calls) for (d, n) if one was built; pf = slabs per step (1 or 2)
template<typename T>
bool half_kernel(int d, int n, int pf, KernelInfo *out);

// kron_slab_fwd (exact, forward mode order) for (d, n) if one was built
template<typename T>
bool slab_fwd_kernel(int d, int n, KernelInfo *out);

// kron_pair_fwd + kron_sub_fwd (exact, forward mode order) for (d, n) if built
template<typename T>
bool twopass_fwd_kernels(int d, int n, KernelInfo *pair, KernelInfo *sub);

// kron_sub + kron_pair for (d, n) if built
template<typename T>
bool twopass_kernels(int d, int n, bool exact, KernelInfo *sub,
                     KernelInfo *pair);

// kron_xcd (km_xcd.hip): the two-pass family fused into one cooperative
// launch, W in each XCD's L2 (FMA calls), if built for (d, n)
template<typename T>
bool xcd_kernel(int d, int n, KernelInfo *k);
// pick the next control block of the current device and queue its reset on
// st; the kernel takes *ctl
hipError_t xcd_prepare(hipStream_t st, unsigned **ctl);
// sticky count of kron_xcd launches whose bounded waits gave up (current
// device; synchronous)
hipError_t xcd_faults(unsigned long long *count);

// matrix-core pass 1 of the two-pass family (km_mixed.hip) for precision
// KM_PRECISION_FP64_MFMA or KM_PRECISION_F16X4, if built for (d, n)
template<typename T>
bool twopass_mx_kernel(int d, int n, int precision, KernelInfo *sub);

// kron_global (always available)
template<typename T>
KernelInfo global_kernel(bool exact);

// ---------------------------------------------------------------------------
// device-built output order (km_order.hip)
// ---------------------------------------------------------------------------
// the builder sorts in workgroups of 4096 entries; at most this many
constexpr int kOrderMaxBlocks = 1024;
constexpr long long kOrderMaxBatch = 4096LL * kOrderMaxBlocks;
// and cut the order into at most this many workgroup ranges
constexpr int kOrderMaxGroups = 8192;

struct OrderOut
{
  int *order;  // batch entry indices, grouped by output, entry order kept
  int *ranges; // G + 1 range starts: the non-empty ranges first, then
               // empty ones at the end of the order
  int *nout;   // [0] number of distinct outputs, [1] non-empty ranges
};
// bytes of device buffer order_build needs
size_t order_bytes(long long batch, int G);
// build the order of pY into buf (device memory of order_bytes), async on st
hipError_t order_build(const void *const *pY, int batch, int G, void *buf,
                       hipStream_t st, OrderOut *out);
// *flag = gen when pY differs anywhere from ybind (batch pointers), async
hipError_t plan_check_launch(const void *const *pY, const void *const *ybind,
                             int batch, int *flag, unsigned gen, hipStream_t st);
// ---------------------------------------------------------------------------
// kernel-variant table (km_tune.cpp)
// ---------------------------------------------------------------------------
// floor(log2 batch): the batch-size bin of a tuned entry
int batch_bin(long long batch);
// variant (km_kernel_variant) recorded for an FMA call, KM_KERNEL_DEFAULT if none
int tuned_variant(int d, int n, int elem_bytes, int owned, long long batch);
// the kernel of variant v for square (d, n) FMA calls, if built
template<typename T>
bool variant_kernel(int v, int d, int n, KernelInfo *k);

// KM_ERROR_HIP (with km_last_error text) when two HIP runtime libraries are
// loaded in the process, else KM_SUCCESS (km_api.cpp)
km_status runtime_check();

// device address of the allocation-free entry point's static buffer on the
// current device (KM_NOALLOC_SCRATCH_BYTES bytes; km_noalloc.hip)
hipError_t noalloc_scratch(void **ptr);

} // namespace km

// Instantiation helper for one (T, D, N): fills *out with the variant asked for.
#define KM_LDS_CASE(T, D, N)                                                   \
  case N:                                                                      \
  {                                                                            \
    constexpr int NT_ = ::km::lds_threads(::km::cpow(N, D));                   \
    if (exact && fwd)                                                          \
      out->fn = (const void *)&::km::kron_lds<T, D, N, NT_, true, true>;       \
    else if (exact)                                                            \
      out->fn = (const void *)&::km::kron_lds<T, D, N, NT_, true, false>;      \
    else                                                                       \
      out->fn = (const void *)&::km::kron_lds<T, D, N, NT_, false, false>;     \
    out->nt = NT_;                                                             \
    out->family = ::km::FAM_LDS;                                               \
    return true;                                                               \
  }
