// Device-wide exclusive scan (reduce-then-scan), parameterised by load/store functors so
// that producers and consumers fuse into the two passes: a Load functor yields element
// i's value, a Store functor receives (i, exclusive prefix, value). Tiles are 256
// threads x 8 consecutive items per thread for large inputs, 1 item below 2M elements.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "svo_device.h"

namespace svo {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;  // items per thread for large inputs
constexpr uint64_t kScanTile = (uint64_t)kScanBlock * kScanItems;
// below this many elements each thread takes one item, so that the grid still has enough
// blocks to fill the chip
constexpr uint64_t kScanSmall = 1ull << 21;

// n_dev (optional) holds the actual element count on the device; n is then an upper
// bound used only to size the grid, and elements at or beyond *n_dev are treated as
// absent (loads give 0, stores are skipped). Each thread owns ITEMS consecutive
// elements: all its loads are issued before the single block-wide scan.
template <typename T, typename Load, int ITEMS>
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(Load ld, uint64_t n,
                                                            const unsigned long long* n_dev,
                                                            T* sums) {
  __shared__ T sm[kScanBlock / kWave];
  if (n_dev) n = min(n, (uint64_t)*n_dev);
  const uint64_t tile = (uint64_t)kScanBlock * ITEMS;
  const uint64_t base = (uint64_t)blockIdx.x * tile + (uint64_t)threadIdx.x * ITEMS;
  T v[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; k++) v[k] = (base + k < n) ? ld(base + k) : (T)0;
  T acc = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; k++) acc += v[k];
  T tot;
  block_incl_scan<T, kScanBlock>(acc, sm, tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <typename T, typename Load, typename Store, int ITEMS>
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(Load ld, Store st, uint64_t n,
                                                           const unsigned long long* n_dev,
                                                           const T* prefix) {
  __shared__ T sm[kScanBlock / kWave];
  if (n_dev) n = min(n, (uint64_t)*n_dev);
  const uint64_t tile = (uint64_t)kScanBlock * ITEMS;
  if ((uint64_t)blockIdx.x * tile >= n) return;  // uniform per block
  const uint64_t base = (uint64_t)blockIdx.x * tile + (uint64_t)threadIdx.x * ITEMS;
  T v[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; k++) v[k] = (base + k < n) ? ld(base + k) : (T)0;
  T acc = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; k++) acc += v[k];
  T tot;
  const T inc = block_incl_scan<T, kScanBlock>(acc, sm, tot);
  T run = (prefix ? prefix[blockIdx.x] : (T)0) + inc - acc;
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    if (base + k < n) st(base + k, run, v[k]);
    run += v[k];
  }
}

template <typename T>
struct ArrayLoad {
  const T* a;
  __device__ T operator()(uint64_t i) const { return a[i]; }
};
template <typename T>
struct ArrayStoreExcl {
  T* a;
  __device__ void operator()(uint64_t i, T ex, T) const { a[i] = ex; }
};

// Number of T elements of scratch that scan_exclusive needs for n elements (sized for
// the one-item tiles, which need the most).
inline uint64_t scan_scratch_elems(uint64_t n) {
  uint64_t tot = 0;
  while (n > (uint64_t)kScanBlock) {
    n = (n + kScanBlock - 1) / kScanBlock;
    tot += n;
  }
  return tot + 1;
}

template <typename T, int ITEMS, typename Load, typename Store>
hipError_t scan_exclusive_items(Load ld, Store st, uint64_t n, T* scratch, hipStream_t s,
                                const unsigned long long* n_dev);

// Exclusive scan of n elements ('n_dev' optional: actual count on the device, n being an
// upper bound). 'scratch' holds scan_scratch_elems(n) T's. Stores for an empty input
// (actual count 0) never run, so a Store that publishes the total must have its total
// slot cleared beforehand.
template <typename T, typename Load, typename Store>
hipError_t scan_exclusive(Load ld, Store st, uint64_t n, T* scratch, hipStream_t s,
                          const unsigned long long* n_dev = nullptr) {
  if (n >= kScanSmall) return scan_exclusive_items<T, kScanItems>(ld, st, n, scratch, s, n_dev);
  return scan_exclusive_items<T, 1>(ld, st, n, scratch, s, n_dev);
}

template <typename T, int ITEMS, typename Load, typename Store>
hipError_t scan_exclusive_items(Load ld, Store st, uint64_t n, T* scratch, hipStream_t s,
                                const unsigned long long* n_dev) {
  if (n == 0) return hipSuccess;
  const uint64_t tile = (uint64_t)kScanBlock * ITEMS;
  const uint64_t nb = (n + tile - 1) / tile;
  if (nb == 1) {
    hipLaunchKernelGGL((k_scan_apply<T, Load, Store, ITEMS>), dim3(1), dim3(kScanBlock), 0, s, ld,
                       st, n, n_dev, (const T*)nullptr);
    return hipGetLastError();
  }
  T* sums = scratch;
  hipLaunchKernelGGL((k_scan_reduce<T, Load, ITEMS>), dim3((unsigned)nb), dim3(kScanBlock), 0, s,
                     ld, n, n_dev, sums);
  hipError_t e = scan_exclusive<T>(ArrayLoad<T>{sums}, ArrayStoreExcl<T>{sums}, nb,
                                   scratch + nb, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_scan_apply<T, Load, Store, ITEMS>), dim3((unsigned)nb), dim3(kScanBlock),
                     0, s, ld, st, n, n_dev, (const T*)sums);
  return hipGetLastError();
}

}  // namespace svo
