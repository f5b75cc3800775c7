// Device-side building blocks shared by the voxelize, partition and build kernels.
//
// Everything here is integer/bit work or FP32 arithmetic in the exact expression order
// of voxelizer.cpp:189-303 (compiled with -ffp-contract=off, IEEE division/sqrt and
// denormals preserved, so each operation rounds exactly as the reference's x86 SSE
// build does).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svo {

constexpr int kWave = 64;

// ------------------------------------------------------------------------------------
// Morton codes (interleave x -> bit 0, y -> bit 1, z -> bit 2 of each triple), the
// closed form of libmorton's morton3D_64_encode/decode for coordinates below 2^21.
// ------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint64_t spread3(uint64_t v) {
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}
__host__ __device__ __forceinline__ uint32_t compact3(uint64_t v) {
  v &= 0x1249249249249249ull;
  v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ull;
  v = (v ^ (v >> 4)) & 0x100f00f00f00f00full;
  v = (v ^ (v >> 8)) & 0x1f0000ff0000ffull;
  v = (v ^ (v >> 16)) & 0x1f00000000ffffull;
  v = (v ^ (v >> 32)) & 0x1fffffull;
  return (uint32_t)v;
}
__host__ __device__ __forceinline__ uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}
__host__ __device__ __forceinline__ void unmorton3(uint64_t m, uint32_t& x, uint32_t& y,
                                                   uint32_t& z) {
  x = compact3(m);
  y = compact3(m >> 1);
  z = compact3(m >> 2);
}

// ------------------------------------------------------------------------------------
// Scalar semantics of the reference's x86 build
// ------------------------------------------------------------------------------------
// static_cast<int>(float) as cvttss2si: NaN / out of range -> INT_MIN
__device__ __forceinline__ int trunc_i32_x86(float f) {
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) return (int)0x80000000;
  return (int)f;
}
__device__ __forceinline__ float min_std(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float max_std(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ int clamp_i(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// ------------------------------------------------------------------------------------
// Partition hit boxes (partitioner.cpp:43-77,101-149, intersection.h:49-53): the range of
// partition cubes along each axis that a triangle's bbox touches, from the cubes' world
// bounds cube_lo[c], cube_hi[c] (c < K per axis), packed as 8-bit (lo, hi) pairs per
// axis; ~0 when the bbox misses every cube. NaN never rejects.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t part_hitbox_of(const float* __restrict__ p,
                                                   const float* __restrict__ cube_lo,
                                                   const float* __restrict__ cube_hi, int K) {
  uint64_t packed = 0;
  bool none = false;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float mn = min_std(p[a], min_std(p[3 + a], p[6 + a]));
    const float mx = max_std(p[a], max_std(p[3 + a], p[6 + a]));
    // rejected on this axis: mx < lo_c or mn > hi_c (NaN never rejects)
    int below = 0, above = 0;
    for (int c = 0; c < K; c++) {
      below += (cube_hi[c] < mn) ? 1 : 0;
      above += (mx < cube_lo[c]) ? 1 : 0;
    }
    const int clo = below, chi = K - 1 - above;
    if (clo > chi) none = true;
    packed |= ((uint64_t)(clo & 0xFF) | ((uint64_t)(chi & 0xFF) << 8)) << (16 * a);
  }
  return none ? ~0ull : packed;
}

// 1 when the triangle with hit box hb[i] touches partition cube (cx, cy, cz)
struct PartLoad {
  const uint64_t* hb;
  uint32_t cx, cy, cz;
  __device__ uint32_t operator()(uint64_t i) const {
    const uint64_t h = hb[i];
    if (h == ~0ull) return 0u;
    const uint32_t c[3] = {cx, cy, cz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const uint32_t lo = (uint32_t)(h >> (16 * a)) & 0xFF, hi = (uint32_t)(h >> (16 * a + 8)) & 0xFF;
      if (c[a] < lo || c[a] > hi) return 0u;
    }
    return 1u;
  }
};

// NaN written to .octreedata is stored as the x86 default NaN 0xFFC00000
__device__ __forceinline__ float canon_nan(float f) {
  return (f != f) ? __uint_as_float(0xFFC00000u) : f;
}

// ------------------------------------------------------------------------------------
// Wave-level helpers (64 lanes)
// ------------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }
// Order a wave's LDS stores before its lanes read one another's values (and those reads
// before the wave's next stores): a wave's LDS operations complete in issue order, so a
// wavefront-scope fence (which keeps the compiler from moving them) and a wave barrier
// suffice.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ T shfl_up_t(T v, int d);
template <>
__device__ __forceinline__ uint32_t shfl_up_t<uint32_t>(uint32_t v, int d) {
  return (uint32_t)__shfl_up((int)v, d, kWave);
}
template <>
__device__ __forceinline__ uint64_t shfl_up_t<uint64_t>(uint64_t v, int d) {
  int lo = __shfl_up((int)(uint32_t)v, d, kWave);
  int hi = __shfl_up((int)(uint32_t)(v >> 32), d, kWave);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
template <typename T>
__device__ __forceinline__ T shfl_t(T v, int src);
template <>
__device__ __forceinline__ uint32_t shfl_t<uint32_t>(uint32_t v, int s) {
  return (uint32_t)__shfl((int)v, s, kWave);
}
template <>
__device__ __forceinline__ uint64_t shfl_t<uint64_t>(uint64_t v, int s) {
  int lo = __shfl((int)(uint32_t)v, s, kWave);
  int hi = __shfl((int)(uint32_t)(v >> 32), s, kWave);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// inclusive scan over the 64 lanes of a wave (all lanes must be active)
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T o = shfl_up_t<T>(v, d);
    if (l >= d) v += o;
  }
  return v;
}
// 32-bit form with data-parallel moves (DPP) instead of LDS permutes: shifts by 1, 2, 4
// and 8 lanes within each row of 16 (lanes shifted in from outside the row read 0),
// then lane 15 of rows 0 and 2 is added to rows 1 and 3, and lane 31 to rows 2 and 3
// (rows the row mask leaves out add 0).
template <>
__device__ __forceinline__ uint32_t wave_incl_scan<uint32_t>(uint32_t v) {
  int x = (int)v;
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return (uint32_t)x;
}
// 64-bit form: the same moves on both halves of the value, 64-bit adds
template <int CTRL, int ROWS, bool BOUND>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, ROWS, 0xf, BOUND);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, ROWS, 0xf, BOUND);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
template <>
__device__ __forceinline__ uint64_t wave_incl_scan<uint64_t>(uint64_t v) {
  v += dpp_u64<0x111, 0xf, true>(v);
  v += dpp_u64<0x112, 0xf, true>(v);
  v += dpp_u64<0x114, 0xf, true>(v);
  v += dpp_u64<0x118, 0xf, true>(v);
  v += dpp_u64<0x142, 0xa, false>(v);
  v += dpp_u64<0x143, 0xc, false>(v);
  return v;
}

// inclusive scan over a block of BLOCK threads; 'total' receives the block sum.
// smem must hold BLOCK/64 elements. All threads of the block must call it.
template <typename T, int BLOCK>
__device__ __forceinline__ T block_incl_scan(T v, T* smem, T& total) {
  constexpr int NW = BLOCK / kWave;
  const int w = threadIdx.x / kWave;
  const int l = threadIdx.x % kWave;
  T inc = wave_incl_scan<T>(v);
  if (l == kWave - 1) smem[w] = inc;
  __syncthreads();
  T before = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; i++) {
    T s = smem[i];
    if (i < w) before += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return before + inc;
}

// ------------------------------------------------------------------------------------
// Shape of the subtree of one occupied 16^3 brick (an aligned subtree of height 4) from
// its 64 occupancy words, lane l holding word l (voxels 64l..64l+63 in Morton order):
// height 1 = a byte of a word, 2 = a word, 3 = 8 consecutive words, 4 = the brick.
// Sizes count the subtree's nodes: s2 = 1 + non-zero bytes + voxels, s3 = 1 + sum s2,
// s4 = 1 + sum s3 (non-null children only). All lanes of the wave must call it.
// ------------------------------------------------------------------------------------
// sum over the 8 lanes of a lane's group
// (DPP moves: lane ^ 1 and lane ^ 2 by quad permutations, lane ^ 4 by a 4-lane shift
// left or right within the 16-lane row)
__device__ __forceinline__ uint32_t group8_sum(uint32_t v) {
  int x = (int)v;
  x += __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xf, 0xf, true);  // quad_perm [1,0,3,2]
  x += __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xf, 0xf, true);  // quad_perm [2,3,0,1]
  const int up = __builtin_amdgcn_update_dpp(0, x, 0x104, 0xf, 0xf, true);    // row_shl:4
  const int down = __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
  x += (lane_id() & 4) ? down : up;
  return (uint32_t)x;
}

struct BrickShape {
  uint64_t word;
  uint32_t m2;      // non-zero bytes of this word
  uint32_t s2;      // subtree size of this word's node (0 if empty)
  uint32_t mask3;   // non-zero words of this lane's group
  uint32_t s3;      // subtree size of the group's node (0 if empty)
  uint32_t mask4;   // non-zero groups
  uint32_t s4;      // subtree size of the brick node
};

__device__ __forceinline__ BrickShape brick_shape_of(uint64_t word) {
  BrickShape B;
  const int lane = lane_id();
  B.word = word;
  B.m2 = 0;
#pragma unroll
  for (int j = 0; j < 8; j++)
    if ((B.word >> (8 * j)) & 0xFFull) B.m2 |= 1u << j;
  B.s2 = B.word ? 1u + (uint32_t)__popc(B.m2) + (uint32_t)__popcll(B.word) : 0u;
  const uint64_t nz = __ballot(B.word != 0ull);
  const int g = lane >> 3;
  B.mask3 = (uint32_t)(nz >> (8 * g)) & 0xFFu;
  const uint32_t sum2 = group8_sum(B.s2);
  B.s3 = B.mask3 ? 1u + sum2 : 0u;
  B.mask4 = 0;
#pragma unroll
  for (int q = 0; q < 8; q++)
    if ((nz >> (8 * q)) & 0xFFull) B.mask4 |= 1u << q;
  // s4 = 1 + sum of the groups' s3: each group counted once (its first lane)
  const uint32_t v = wave_incl_scan<uint32_t>(((lane & 7) == 0) ? B.s3 : 0u);
  B.s4 = 1u + (uint32_t)__builtin_amdgcn_readlane((int)v, kWave - 1);
  return B;
}
// s4 (below 2^13) and the brick's child mask in one word
__device__ __forceinline__ uint32_t brick_shape_word(const BrickShape& B) {
  return B.s4 | (B.mask4 << 16);
}

}  // namespace svo
