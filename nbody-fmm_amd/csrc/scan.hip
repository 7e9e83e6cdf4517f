// scan.hip -- device-wide exclusive prefix sums (reduce-then-scan, recursive
// over block totals).  Tiles of 1024 elements: 256 threads x 4 items.
#include "nbfmm_internal.h"

namespace nbfmm {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 4;
constexpr int kTile = kThreads * kItems;

template <class T>
__device__ T block_excl_scan(T v, T *lds, T *total) {
  // wave-level inclusive scan with shuffles, then across the 4 waves via LDS
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  T wsum = 0, all = 0;
  for (int w = 0; w < kThreads / 64; ++w) {
    if (w < wave) wsum += lds[w];
    all += lds[w];
  }
  __syncthreads();
  *total = all;
  return wsum + x - v;
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_scan_tiles(const T *in, T *out, int64_t n, T *sums) {
  __shared__ T lds[kThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  T v[kItems];
  T local = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : T(0);
    local += v[k];
  }
  T total;
  T run = block_excl_scan<T>(local, lds, &total);
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0 && sums) sums[blockIdx.x] = total;
}

template <class T>
__global__ void k_add_offsets(T *out, int64_t n, const T *offs) {
  const int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
  const T o = offs[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    int64_t j = i + (int64_t)k * kThreads;
    if (j < n) out[j] += o;
  }
}

template <class T>
void scan_rec(const T *in, T *out, int64_t n, T *tmp, hipStream_t s) {
  if (n <= 0) return;
  int64_t nb = (n + kTile - 1) / kTile;
  T *sums = nb > 1 ? tmp : nullptr;
  hipLaunchKernelGGL(k_scan_tiles<T>, dim3((unsigned)nb), dim3(kThreads), 0, s, in, out, n, sums);
  NB_LAUNCH();
  if (nb > 1) {
    scan_rec<T>(sums, sums, nb, tmp + nb, s);
    hipLaunchKernelGGL(k_add_offsets<T>, dim3((unsigned)nb), dim3(kThreads), 0, s, out, n,
                       (const T *)sums);
    NB_LAUNCH();
  }
}

template <class T>
void scan_excl(const T *in, T *out, int64_t n, DBuf &tmp, hipStream_t s) {
  int64_t need = 0, m = n;
  while (m > kTile) {
    m = (m + kTile - 1) / kTile;
    need += m + 1;
  }
  T *t = dget<T>(tmp, (size_t)need + 64);
  scan_rec<T>(in, out, n, t, s);
}

} // namespace

void scan_excl_i32(const int32_t *in, int32_t *out, int64_t n, DBuf &tmp, hipStream_t s) {
  scan_excl<int32_t>(in, out, n, tmp, s);
}
void scan_excl_i64(const int64_t *in, int64_t *out, int64_t n, DBuf &tmp, hipStream_t s) {
  scan_excl<int64_t>(in, out, n, tmp, s);
}
void scan_excl_f64(const double *in, double *out, int64_t n, DBuf &tmp, hipStream_t s) {
  scan_excl<double>(in, out, n, tmp, s);
}

} // namespace nbfmm
