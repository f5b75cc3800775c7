// Device-wide exclusive scan (reduce-then-scan), parameterised by load/store functors so
// that producers and consumers fuse into the two passes: a Load functor yields element
// i's value, a Store functor receives (i, exclusive prefix, value, i is the last element:
// the count is known to the kernel, so a store need not read it back). Tiles are 256
// threads x 8 items per thread (wave-striped) for large inputs, 1 item below 2M elements.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "svo_device.h"
#include "svo_internal.h"

namespace svo {

// Tiles are laid out wave-striped: wave w of a block owns the tile's elements
// [seg, seg + 64 ITEMS), seg = tile start + 64 ITEMS w, and its lane l takes elements
// seg + 64 k + l (k < ITEMS), so every load and store instruction of the wave touches 64
// consecutive elements. Within the segment the order is k-major, so a lane's exclusive
// prefix of item k is (rows before k) + (lanes before l in row k).
template <typename T>
__device__ __forceinline__ T wave_last(T v);
template <>
__device__ __forceinline__ uint32_t wave_last<uint32_t>(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, kWave - 1);
}
template <>
__device__ __forceinline__ uint64_t wave_last<uint64_t>(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, kWave - 1);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), kWave - 1);
  return ((uint64_t)hi << 32) | lo;
}
// First element of the lane's striped items (item k is at the result + 64 k).
template <int ITEMS>
__device__ __forceinline__ uint64_t striped_first(uint64_t tile, int block) {
  const uint64_t w = threadIdx.x / kWave;
  return (tile * (uint64_t)block + w * kWave) * ITEMS + lane_id();
}
// Exclusive prefixes (within the block's tile) of the lane's striped items; 'total'
// receives the tile sum. smem holds BLOCK/64 elements; all threads must call it. One
// barrier: smem is not read again after it.
template <typename T, int BLOCK, int ITEMS>
__device__ __forceinline__ void tile_excl_striped(const T (&v)[ITEMS], T (&ex)[ITEMS], T* smem,
                                                  T& total) {
  constexpr int NW = BLOCK / kWave;
  const int w = threadIdx.x / kWave;
  T run = 0;  // the wave's rows before row k
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    const T inc = wave_incl_scan<T>(v[k]);
    ex[k] = run + inc - v[k];
    run += wave_last<T>(inc);
  }
  if (lane_id() == 0) smem[w] = run;
  __syncthreads();
  T before = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; i++) {
    const T s = smem[i];
    if (i < w) before += s;
    tot += s;
  }
#pragma unroll
  for (int k = 0; k < ITEMS; k++) ex[k] += before;
  total = tot;
}

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;  // items per thread for large inputs
constexpr uint64_t kScanTile = (uint64_t)kScanBlock * kScanItems;
// below this many elements each thread takes one item, so that the grid still has enough
// blocks to fill the chip
constexpr uint64_t kScanSmall = 1ull << 21;

// n_dev (optional) holds the actual element count on the device; n is then an upper
// bound used only to size the grid, and elements at or beyond *n_dev are treated as
// absent (loads give 0, stores are skipped). Each thread owns ITEMS wave-striped
// elements: all its loads are issued before the block-wide scan.
template <typename T, typename Load, int ITEMS>
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(Load ld, uint64_t n,
                                                            const unsigned long long* n_dev,
                                                            T* sums) {
  __shared__ T sm[kScanBlock / kWave];
  if (n_dev) n = min(n, (uint64_t)*n_dev);
  const uint64_t first = striped_first<ITEMS>(blockIdx.x, kScanBlock);
  T acc = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    const uint64_t i = first + (uint64_t)kWave * k;
    if (i < n) acc += ld(i);
  }
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
  const uint64_t first = striped_first<ITEMS>(blockIdx.x, kScanBlock);
  T v[ITEMS], ex[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    const uint64_t i = first + (uint64_t)kWave * k;
    v[k] = (i < n) ? ld(i) : (T)0;
  }
  T tot;
  tile_excl_striped<T, kScanBlock, ITEMS>(v, ex, sm, tot);
  const T p = prefix ? prefix[blockIdx.x] : (T)0;
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    const uint64_t i = first + (uint64_t)kWave * k;
    if (i < n) st(i, p + ex[k], v[k], i == n - 1);
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
  __device__ void operator()(uint64_t i, T ex, T, bool) const { a[i] = ex; }
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
    SVO_LAUNCH((k_scan_apply<T, Load, Store, ITEMS>), dim3(1), dim3(kScanBlock), 0, s, ld,
                       st, n, n_dev, (const T*)nullptr);
    return hipGetLastError();
  }
  T* sums = scratch;
  SVO_LAUNCH((k_scan_reduce<T, Load, ITEMS>), dim3((unsigned)nb), dim3(kScanBlock), 0, s,
                     ld, n, n_dev, sums);
  hipError_t e = scan_exclusive<T>(ArrayLoad<T>{sums}, ArrayStoreExcl<T>{sums}, nb,
                                   scratch + nb, s);
  if (e != hipSuccess) return e;
  SVO_LAUNCH((k_scan_apply<T, Load, Store, ITEMS>), dim3((unsigned)nb), dim3(kScanBlock),
                     0, s, ld, st, n, n_dev, (const T*)sums);
  return hipGetLastError();
}

// ---------------------------------------------------------------- single-pass scan
// One launch (decoupled look-back, Merrill & Garland 2016): a block takes the next tile
// number from a ticket counter, scans that tile, publishes its tile's sum, then its wave 0
// walks back over the predecessors' status words until it meets an inclusive prefix,
// publishes its own inclusive prefix and the block applies the stores. A status word is
// the flag (bits 63..62: 0 = not yet published, 1 = the tile's own sum, 2 = inclusive
// prefix through the tile) and the value (bits 61..0; every prefix must stay below 2^62)
// in one word, written and read with device-scope atomics: those bypass the per-XCD L2
// caches, so the word is coherent across the chip without a fence. With tiles numbered in
// the order blocks start, every tile a block waits on belongs to a block that is already
// running. (Numbering by block index is not enough: the chip hands workgroups to its
// eight XCDs in turn, and a block whose XCD is full waits while its successor runs on
// another; with a second stream's look-back blocks filling the XCDs, two scans could wait
// on each other.) The wait is bounded all the same: after kScanSpinLimit polls the block
// counts an error in *err and goes on with a zero prefix.
//
// Blocks are 1024 threads (tiles of 8192 elements for large inputs), so that the walk
// back is short; tiles at or beyond the actual element count exit at once (no later tile
// waits on them).
//
// status: 1 + tiles words, zero before the launch ([0] is the ticket counter).
constexpr int kSingleBlock = 1024;
constexpr uint64_t kSingleSmall = 1ull << 16;  // below: one item per thread
constexpr unsigned long long kStAgg = 1ull << 62;
constexpr unsigned long long kStIncl = 2ull << 62;
constexpr unsigned long long kStVal = (1ull << 62) - 1;
constexpr uint32_t kScanSpinLimit = 1u << 22;
constexpr int kLookWindows = 4;  // 64-tile windows read per round of the walk back

__device__ __forceinline__ void status_put(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long status_get(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T, typename Load, typename Store, int ITEMS>
__global__ __launch_bounds__(kSingleBlock) void k_scan_single(Load ld, Store st, uint64_t n,
                                                              const unsigned long long* n_dev,
                                                              unsigned long long* status,
                                                              unsigned long long* err) {
  __shared__ T sm[kSingleBlock / kWave];
  __shared__ T s_prefix;
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = (uint32_t)atomicAdd(status, 1ull);
  __syncthreads();
  const uint32_t tile = s_tile;
  unsigned long long* tiles = status + 1;
  if (n_dev) n = min(n, (uint64_t)*n_dev);
  if ((uint64_t)tile * kSingleBlock * ITEMS >= n) return;  // uniform per block
  const uint64_t first = striped_first<ITEMS>(tile, kSingleBlock);
  T v[ITEMS], ex[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    const uint64_t i = first + (uint64_t)kWave * k;
    v[k] = (i < n) ? ld(i) : (T)0;
  }
  T tot;
  tile_excl_striped<T, kSingleBlock, ITEMS>(v, ex, sm, tot);
  if (threadIdx.x < kWave) {  // wave 0: publish, look back
    T prefix = 0;
    if (tile == 0) {
      if (threadIdx.x == 0) status_put(&tiles[0], kStIncl | (unsigned long long)tot);
    } else {
      if (threadIdx.x == 0) status_put(&tiles[tile], kStAgg | (unsigned long long)tot);
      // lane l looks at tiles j - l - 64 r, r < kLookWindows: the windows' status words
      // are read together, so a walk back over up to 256 tiles costs one round trip
      int64_t j = (int64_t)tile - 1;
      uint32_t polls = 0;
      bool found = false;
      while (!found) {
        unsigned long long s[kLookWindows];
#pragma unroll
        for (int r = 0; r < kLookWindows; r++) {
          const int64_t idx = j - (int64_t)lane_id() - (int64_t)kWave * r;
          s[r] = (idx >= 0) ? status_get(&tiles[idx]) : kStIncl;
        }
#pragma unroll
        for (int r = 0; r < kLookWindows; r++) {
          if (found) break;
          const int64_t idx = j - (int64_t)lane_id() - (int64_t)kWave * r;
          unsigned long long w = s[r];
          for (;;) {
            const bool pend = (w >> 62) == 0ull;
            if (!__ballot(pend)) break;
            if (++polls > kScanSpinLimit) {
              if (lane_id() == 0) atomicAdd(err, 1ull);
              if (pend) w = kStIncl;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            if (pend) w = status_get(&tiles[idx]);
          }
          const unsigned long long incl = __ballot((w >> 62) == 2ull);
          const int first = incl ? __ffsll((long long)incl) - 1 : kWave;
          const T val = (lane_id() <= first) ? (T)(w & kStVal) : (T)0;
          prefix += wave_last<T>(wave_incl_scan<T>(val));
          found = incl != 0ull;
        }
        j -= (int64_t)kWave * kLookWindows;
      }
      if (threadIdx.x == 0) status_put(&tiles[tile], kStIncl | (unsigned long long)(prefix + tot));
    }
    if (threadIdx.x == 0) s_prefix = prefix;
  }
  __syncthreads();
  const T p = s_prefix;
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    const uint64_t i = first + (uint64_t)kWave * k;
    if (i < n) st(i, p + ex[k], v[k], i == n - 1);
  }
}

// Items per thread of the single-pass scan for n elements, and its status words.
inline int scan_single_items(uint64_t n) { return n >= kSingleSmall ? kScanItems : 1; }
inline uint64_t scan_single_words(uint64_t n) {
  const uint64_t tile = (uint64_t)kSingleBlock * scan_single_items(n);
  return 1 + (n + tile - 1) / tile;
}

// Exclusive scan in one launch; 'status' holds scan_single_words(n) zero words. n is an
// upper bound when n_dev is given (as for scan_exclusive).
template <typename T, typename Load, typename Store>
hipError_t scan_single(Load ld, Store st, uint64_t n, unsigned long long* status,
                       unsigned long long* err, hipStream_t s,
                       const unsigned long long* n_dev = nullptr) {
  if (n == 0) return hipSuccess;
  const int items = scan_single_items(n);
  const uint64_t tile = (uint64_t)kSingleBlock * items;
  const unsigned nb = (unsigned)((n + tile - 1) / tile);
  if (items == kScanItems)
    SVO_LAUNCH((k_scan_single<T, Load, Store, kScanItems>), dim3(nb), dim3(kSingleBlock),
                       0, s, ld, st, n, n_dev, status, err);
  else
    SVO_LAUNCH((k_scan_single<T, Load, Store, 1>), dim3(nb), dim3(kSingleBlock), 0, s,
                       ld, st, n, n_dev, status, err);
  return hipGetLastError();
}

}  // namespace svo
