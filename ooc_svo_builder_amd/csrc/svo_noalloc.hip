// svo_partition_noalloc: the device work of svo_partition (partitioner.cpp:43-77,101-149)
// for callers that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc). Its working memory is one file-scope device
// array, `scratch`, of fixed size; the caller passes the triangle records, the list and
// count arrays, and a stream. Nothing here calls hipMalloc, hipFree, hipMemset* or an
// allocating library routine, and nothing synchronizes: the results are ready when the
// caller's stream reaches them.
//
// scratch (per device, held until the process ends), for up to kNaMaxTris triangles and
// kNaMaxParts partitions:
//   hit boxes     8 B per triangle (packed partition-cube ranges)
//   cube bounds   2 x 16 floats (world bounds of the cubes along an axis)
//   run           u64 per partition + 1: first list entry of each partition
//   status        look-back words of the order-preserving scan of one partition
//   err           scan look-back time-outs
// Every call on a device shares that device's array, so calls that overlap on one device
// (two streams, two threads) race; run them one after another.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/svo_amd.h"
#include "svo_device.h"
#include "svo_internal.h"
#include "svo_scan.h"

namespace svo {

namespace {
constexpr uint64_t kNaMaxTris = SVO_NOALLOC_MAX_TRIANGLES;
constexpr uint64_t kNaMaxParts = SVO_NOALLOC_MAX_PARTITIONS;
constexpr int kNaMaxK = 16;  // cubes per axis: kNaMaxParts = 16^3
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaHitOff = 0;
constexpr size_t kNaBoundsOff = align256(kNaHitOff + kNaMaxTris * 8);
constexpr size_t kNaRunOff = align256(kNaBoundsOff + 2 * kNaMaxK * sizeof(float));
constexpr size_t kNaErrOff = align256(kNaRunOff + (kNaMaxParts + 1) * 8);
constexpr size_t kNaStatusOff = align256(kNaErrOff + 8);
// look-back words of one scan of kNaMaxTris elements (at least one element per thread)
constexpr size_t kNaStatusWords = 1 + (kNaMaxTris + kSingleBlock - 1) / kSingleBlock;
constexpr size_t kNaScratchBytes = align256(kNaStatusOff + kNaStatusWords * 8);
}  // namespace

static __device__ __align__(256) uint8_t scratch[kNaScratchBytes];

// Device address of this device's `scratch`, looked up once per device with that device
// current (the name itself is a host-side stand-in, not a device address).
static std::mutex g_na_mu;
static void* g_na_ptr[64] = {};

static hipError_t scratch_on(int dev, uint8_t** out) {
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_na_ptr[dev] = p;
  }
  *out = (uint8_t*)g_na_ptr[dev];
  return hipSuccess;
}

// run[0..P] = 0, err = 0, and the cubes' world bounds along an axis, with the host
// path's expression (svo_voxelize.hip partition_job): (float)(uint32_t)(c * side) * unit
__global__ __launch_bounds__(256) void k_na_init(unsigned long long* run, uint64_t n_run,
                                                 unsigned long long* err, float* lo, float* hi,
                                                 int K, uint64_t side, float unit) {
  for (uint64_t i = threadIdx.x; i < n_run; i += blockDim.x) run[i] = 0;
  if (threadIdx.x == 0) *err = 0;
  if (threadIdx.x < (unsigned)K) {
    const uint64_t c = threadIdx.x;
    lo[c] = (float)(uint32_t)(c * side) * unit;
    hi[c] = (float)(uint32_t)((c + 1) * side) * unit;
  }
}

__global__ __launch_bounds__(256) void k_na_hitbox(const float* __restrict__ tris, int stride,
                                                   uint64_t n, const float* __restrict__ lo,
                                                   const float* __restrict__ hi, int K,
                                                   uint64_t* __restrict__ hitbox) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) hitbox[t] = part_hitbox_of(tris + t * stride, lo, hi, K);
}

__global__ __launch_bounds__(256) void k_na_zero(unsigned long long* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// One partition: list entries from run[j] on (those beyond the capacity are dropped),
// its count, and where the next partition's list starts.
struct NaStore {
  uint32_t* lists;
  uint64_t cap;
  unsigned long long* run;
  uint64_t j;
  uint64_t* counts;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t v, bool last) const {
    const uint64_t base = run[j];
    if (v && base + ex < cap) lists[base + ex] = (uint32_t)i;
    if (last) {
      counts[j] = (uint64_t)ex + v;
      run[j + 1] = base + ex + v;
    }
  }
};

// One partition (svo_partition's partition_one case): every triangle, in order.
__global__ __launch_bounds__(256) void k_na_identity(uint32_t* lists, uint64_t cap, uint64_t n,
                                                     uint64_t* counts, unsigned long long* run) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && i < cap) lists[i] = (uint32_t)i;
  if (i == 0) {
    counts[0] = n;
    run[1] = n;
  }
}

// counts[P]: 0 when every list is complete, 1 when the lists exceed the capacity, 2 when
// a scan's look-back wait ran out
__global__ void k_na_finish(const unsigned long long* run, uint64_t P, uint64_t cap,
                            const unsigned long long* err, uint64_t* counts) {
  if (threadIdx.x == 0) counts[P] = *err ? 2 : (run[P] > cap ? 1 : 0);
}

}  // namespace svo

using namespace svo;

extern "C" int svo_partition_noalloc(void* hip_stream, const float* d_records,
                                     uint64_t n_triangles, int payload, float unitlength,
                                     uint64_t gridsize, uint64_t n_partitions, uint32_t* d_lists,
                                     uint64_t lists_capacity, uint64_t* d_counts) {
  set_error("");
  const hipStream_t s = (hipStream_t)hip_stream;
  if (n_triangles > kNaMaxTris) {
    set_error("svo_partition_noalloc: " + std::to_string(n_triangles) +
              " triangles, at most " + std::to_string(kNaMaxTris));
    return SVO_ERR_INVALID;
  }
  if (gridsize == 0 || (gridsize & (gridsize - 1)) || n_partitions == 0 ||
      n_partitions > kNaMaxParts || !d_counts || (n_triangles && (!d_records || !d_lists))) {
    set_error("svo_partition_noalloc: invalid arguments (gridsize a power of 2, "
              "1 <= n_partitions <= " + std::to_string(kNaMaxParts) + ", device arrays)");
    return SVO_ERR_INVALID;
  }
  int k = 0;
  while ((1ull << (3 * k)) < n_partitions) k++;
  const int K = 1 << k;
  if ((1ull << (3 * k)) != n_partitions || (uint64_t)K > gridsize) {
    set_error("svo_partition_noalloc: n_partitions must be a power of 8 not above gridsize^3");
    return SVO_ERR_INVALID;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  uint8_t* base = nullptr;
  if (e == hipSuccess) e = scratch_on(dev, &base);
  if (e != hipSuccess) {
    set_error(std::string("svo_partition_noalloc: ") + hipGetErrorString(e));
    return SVO_ERR_DEVICE;
  }
  uint64_t* hitbox = (uint64_t*)(base + kNaHitOff);
  float* lo = (float*)(base + kNaBoundsOff);
  float* hi = lo + kNaMaxK;
  unsigned long long* run = (unsigned long long*)(base + kNaRunOff);
  unsigned long long* err = (unsigned long long*)(base + kNaErrOff);
  unsigned long long* status = (unsigned long long*)(base + kNaStatusOff);
  const uint64_t n = n_triangles;
  const uint64_t side = gridsize / (uint64_t)K;
  SVO_LAUNCH(k_na_init, dim3(1), dim3(256), 0, s, run, n_partitions + 1, err, lo, hi, K,
                     side, unitlength);
  if (n_partitions == 1) {
    SVO_LAUNCH(k_na_identity, dim3((unsigned)std::max<uint64_t>(1, (n + 255) / 256)),
                       dim3(256), 0, s, d_lists, lists_capacity, n, d_counts, run);
  } else {
    const int stride = payload ? 21 : 9;
    if (n)
      SVO_LAUNCH(k_na_hitbox, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                         d_records, stride, n, (const float*)lo, (const float*)hi, K, hitbox);
    const uint64_t words = scan_single_words(n);
    for (uint64_t j = 0; j < n_partitions; j++) {
      uint32_t cx, cy, cz;
      unmorton3(j, cx, cy, cz);
      if (n == 0) {
        // an empty mesh: every count 0 (the last element of a scan writes them otherwise)
        SVO_LAUNCH(k_na_zero, dim3(1), dim3(256), 0, s,
                           (unsigned long long*)d_counts, n_partitions);
        break;
      }
      SVO_LAUNCH(k_na_zero, dim3(1), dim3(256), 0, s, status, words);
      e = scan_single<uint32_t>(PartLoad{hitbox, cx, cy, cz},
                                NaStore{d_lists, lists_capacity, run, j, d_counts}, n, status,
                                err, s);
      if (e != hipSuccess) break;
    }
  }
  if (e == hipSuccess) {
    SVO_LAUNCH(k_na_finish, dim3(1), dim3(64), 0, s, (const unsigned long long*)run,
                       n_partitions, lists_capacity, (const unsigned long long*)err, d_counts);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    set_error(std::string("svo_partition_noalloc: ") + hipGetErrorString(e));
    return SVO_ERR_DEVICE;
  }
  return SVO_OK;
}
