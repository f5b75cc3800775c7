// Partitioner and voxelizer kernels.
//
// Partition (partitioner.cpp:43-77,101-149, BBoxBuffer.h:70-84): per triangle, the box
// of Morton-range partitions whose world box its bbox touches (inclusive test,
// intersection.h:49-53), then per partition an order-preserving flag scan that writes
// the partition's triangle index list (histogram -> scan -> scatter).
//
// Voxelize (voxelizer.cpp:138-307): the Schwarz-Seidel conservative triangle/box test in
// FP32 with the reference's expression order. Work is split in two tiers:
//   * small triangles: one thread walks all columns of its triangle's clamped bbox;
//   * large triangles: one thread per column, the columns of all large triangles laid
//     out by an exclusive scan of their column counts.
// Columns run along the axis of the triangle's dominant normal component. Per column the
// projection test in the plane across that axis is evaluated once (it does not depend
// on the position along the column), and only the voxels within two voxels of the
// plane slab are tested; every voxel outside that range fails the reference's plane
// test (the margin is far above the FP32 error of the test). Triangles with a
// non-finite normal or far-away vertices walk full columns.
// Marked voxels go to a sparse brick grid; compaction of the occupied bricks in Morton
// order yields the sorted unique codes the builder needs (the reference sorts its side
// array, main.cpp:356-359). Payload voxels are resolved to the first triangle in
// partition order that covers them (voxelizer.cpp:263) by a second pass doing
// atomicMin on the triangle index per voxel rank.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>

#include "svo_device.h"
#include "svo_internal.h"
#include "svo_scan.h"

namespace svo {

typedef unsigned long long ull;

struct VoxParams {
  const float* tris;
  const uint32_t* list;  // partition's triangle indices, nullptr = identity
  uint32_t n;
  int stride;  // floats per record: 9 binary, 21 payload
  uint64_t mstart;
  int pmin[3], pmax[3];
  float unit, unit_div;
  float coord_limit;  // |vertex| beyond this walks full columns
  double slab_margin; // voxels added on each side of the exact plane slab
  int small_cols;
};

struct Setup {
  float n[3];
  float d1, d2;
  float ne[9][2];  // edge normals: xy e0..e2, yz e0..e2, zx e0..e2
  float de[9];
  int g0[3], g1[3];
  int w;          // column axis
  bool slab;      // restrict columns to the plane slab
  double nv0, nc, nm;
};

// voxelizer.cpp:189-254 (bbox, clamp, normal, critical point, edge functions)
__device__ __forceinline__ void setup_tri(const float* __restrict__ p, const VoxParams& P,
                                          Setup& S) {
  const float v0x = p[0], v0y = p[1], v0z = p[2];
  const float v1x = p[3], v1y = p[4], v1z = p[5];
  const float v2x = p[6], v2y = p[7], v2z = p[8];
  const float u = P.unit;
  {
    const float mn0 = min_std(v0x, min_std(v1x, v2x)), mx0 = max_std(v0x, max_std(v1x, v2x));
    const float mn1 = min_std(v0y, min_std(v1y, v2y)), mx1 = max_std(v0y, max_std(v1y, v2y));
    const float mn2 = min_std(v0z, min_std(v1z, v2z)), mx2 = max_std(v0z, max_std(v1z, v2z));
    S.g0[0] = clamp_i(trunc_i32_x86(mn0 * P.unit_div), P.pmin[0], P.pmax[0]);
    S.g0[1] = clamp_i(trunc_i32_x86(mn1 * P.unit_div), P.pmin[1], P.pmax[1]);
    S.g0[2] = clamp_i(trunc_i32_x86(mn2 * P.unit_div), P.pmin[2], P.pmax[2]);
    S.g1[0] = clamp_i(trunc_i32_x86(mx0 * P.unit_div), P.pmin[0], P.pmax[0]);
    S.g1[1] = clamp_i(trunc_i32_x86(mx1 * P.unit_div), P.pmin[1], P.pmax[1]);
    S.g1[2] = clamp_i(trunc_i32_x86(mx2 * P.unit_div), P.pmin[2], P.pmax[2]);
  }
  const float e0x = v1x - v0x, e0y = v1y - v0y, e0z = v1z - v0z;
  const float e1x = v2x - v1x, e1y = v2y - v1y, e1z = v2z - v1z;
  const float e2x = v0x - v2x, e2y = v0y - v2y, e2z = v0z - v2z;
  const float cx = e0y * e1z - e1y * e0z;
  const float cy = e0z * e1x - e1z * e0x;
  const float cz = e0x * e1y - e1x * e0y;
  const float ss = (cx * cx + cy * cy) + cz * cz;
  const float inv = 1.0f / sqrtf(ss);
  const float nx = cx * inv, ny = cy * inv, nz = cz * inv;
  S.n[0] = nx;
  S.n[1] = ny;
  S.n[2] = nz;
  const float ccx = (nx > 0) ? u : 0.0f;
  const float ccy = (ny > 0) ? u : 0.0f;
  const float ccz = (nz > 0) ? u : 0.0f;
  S.d1 = (nx * (ccx - v0x) + ny * (ccy - v0y)) + nz * (ccz - v0z);
  S.d2 = (nx * ((u - ccx) - v0x) + ny * ((u - ccy) - v0y)) + nz * ((u - ccz) - v0z);

  const float E[3][3] = {{e0x, e0y, e0z}, {e1x, e1y, e1z}, {e2x, e2y, e2z}};
  const float V[3][3] = {{v0x, v0y, v0z}, {v1x, v1y, v1z}, {v2x, v2y, v2z}};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    // XY plane (voxelizer.cpp:220-230)
    float a = -1.0f * E[k][1], b = E[k][0];
    if (nz < 0.0f) {
      a = -1.0f * a;
      b = -1.0f * b;
    }
    S.ne[k][0] = a;
    S.ne[k][1] = b;
    S.de[k] = ((-1.0f * (a * V[k][0] + b * V[k][1])) + max_std(0.0f, u * a)) + max_std(0.0f, u * b);
    // YZ plane (:232-242)
    a = -1.0f * E[k][2];
    b = E[k][1];
    if (nx < 0.0f) {
      a = -1.0f * a;
      b = -1.0f * b;
    }
    S.ne[3 + k][0] = a;
    S.ne[3 + k][1] = b;
    S.de[3 + k] = ((-1.0f * (a * V[k][1] + b * V[k][2])) + max_std(0.0f, u * a)) + max_std(0.0f, u * b);
    // ZX plane (:244-254)
    a = -1.0f * E[k][0];
    b = E[k][2];
    if (ny < 0.0f) {
      a = -1.0f * a;
      b = -1.0f * b;
    }
    S.ne[6 + k][0] = a;
    S.ne[6 + k][1] = b;
    S.de[6 + k] = ((-1.0f * (a * V[k][2] + b * V[k][0])) + max_std(0.0f, u * a)) + max_std(0.0f, u * b);
  }

  // column axis and slab restriction
  const float ax = fabsf(nx), ay = fabsf(ny), az = fabsf(nz);
  S.w = (az >= ax && az >= ay) ? 2 : (ax >= ay ? 0 : 1);
  const float L = P.coord_limit;
  const bool finite = (ax <= 2.0f) && (ay <= 2.0f) && (az <= 2.0f);  // false for NaN/inf
  const bool near = fabsf(v0x) <= L && fabsf(v0y) <= L && fabsf(v0z) <= L && fabsf(v1x) <= L &&
                    fabsf(v1y) <= L && fabsf(v1z) <= L && fabsf(v2x) <= L && fabsf(v2y) <= L &&
                    fabsf(v2z) <= L;
  S.slab = finite && near && fmaxf(ax, fmaxf(ay, az)) >= 0.5f;
  S.nv0 = ((double)nx * v0x + (double)ny * v0y) + (double)nz * v0z;
  S.nc = ((double)nx * ccx + (double)ny * ccy) + (double)nz * ccz;
  S.nm = ((double)nx * (u - ccx) + (double)ny * (u - ccy)) + (double)nz * (u - ccz);
}

// ---------------------------------------------------------------- column walk
// A triangle's voxels are tested column by column along its column axis w (the axis of
// the dominant normal component); a column is fixed by its coordinates (a, b) on the two
// other axes ua < va. The per-triangle constants are re-slotted once so that one
// branch-free code path serves all three axes:
//   * plane test t = (n0*px + n1*py) + n2*pz becomes t = (nw*pc + Ka) + Kb with
//     Ka, Kb per column (Kb = +0 when w = z; a zero's sign never changes the decision);
//   * the projection test in the plane across w (3 edges) depends on (pa, pb) only and
//     is evaluated once per column: (h0*pa + h1*pb) + hd;
//   * each edge of the two other planes is (cv*pc + cc*pconst) + de per voxel.
// Every product and sum is the reference's own (voxelizer.cpp:257-287); only operands of
// single additions are swapped, which is exact in IEEE arithmetic.
enum WalkF {
  F_NW = 0, F_D1, F_D2, F_NUA, F_NVA,
  F_H0 = 5, F_H1 = 8, F_HD = 11,
  F_AV = 14, F_AC = 17, F_AD = 20,
  F_BV = 23, F_BC = 26, F_BD = 29,
  F_KA = 32, F_KB, F_ALO, F_AHI,
  kWalkF = 36
};
enum WalkI { I_G0U = 0, I_G1U, I_G0V, I_G1V, I_G0W, I_G1W, I_WS, I_TRI, kWalkI = 8 };
constexpr int kWalkWords = kWalkF + kWalkI;

struct Walk {
  float f[kWalkF];
  int i[kWalkI];
  __device__ __forceinline__ int w() const { return i[I_WS] & 3; }
  __device__ __forceinline__ bool slab() const { return (i[I_WS] >> 2) & 1; }
  __device__ __forceinline__ uint64_t columns() const {
    return (uint64_t)(i[I_G1U] - i[I_G0U] + 1) * (uint64_t)(i[I_G1V] - i[I_G0V] + 1);
  }
  __device__ __forceinline__ uint64_t work() const {
    return slab() ? columns() : columns() * (uint64_t)(i[I_G1W] - i[I_G0W] + 1);
  }
};

// Copy one axis' plane roles: HP/AP/BP = first edge of the hoisted plane and of planes
// A and B; HS = the hoisted plane's operands come as (pb, pa); AV0/BV0 = the column
// coordinate is the plane's first operand.
template <int HP, int AP, int BP, bool HS, bool AV0, bool BV0>
__device__ __forceinline__ void fill_planes(const Setup& S, Walk& K) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float hx = S.ne[HP + k][0], hy = S.ne[HP + k][1];
    K.f[F_H0 + k] = HS ? hy : hx;
    K.f[F_H1 + k] = HS ? hx : hy;
    K.f[F_HD + k] = S.de[HP + k];
    const float ax = S.ne[AP + k][0], ay = S.ne[AP + k][1];
    K.f[F_AV + k] = AV0 ? ax : ay;
    K.f[F_AC + k] = AV0 ? ay : ax;
    K.f[F_AD + k] = S.de[AP + k];
    const float bx = S.ne[BP + k][0], by = S.ne[BP + k][1];
    K.f[F_BV + k] = BV0 ? bx : by;
    K.f[F_BC + k] = BV0 ? by : bx;
    K.f[F_BD + k] = S.de[BP + k];
  }
}

// Re-slot a triangle's setup (voxelizer.cpp:189-254, see setup_tri) for the column walk.
__device__ __forceinline__ void make_walk(const Setup& S, const VoxParams& P, uint32_t tri,
                                          Walk& K) {
  const int w = S.w;
  const int ua = (w == 0) ? 1 : 0;
  const int va = (w == 2) ? 1 : 2;
  // edges of the three planes: xy = ne[0..2], yz = ne[3..5], zx = ne[6..8]
  if (w == 2)         // hoisted xy(pa, pb); A = yz(pb, pc); B = zx(pc, pa)
    fill_planes<0, 3, 6, false, false, true>(S, K);
  else if (w == 0)    // hoisted yz(pa, pb); A = xy(pc, pa); B = zx(pb, pc)
    fill_planes<3, 0, 6, false, true, false>(S, K);
  else                // hoisted zx(pb, pa); A = xy(pa, pc); B = yz(pc, pb)
    fill_planes<6, 0, 3, true, false, true>(S, K);
  const float n0 = S.n[0], n1 = S.n[1], n2 = S.n[2];
  const float nw = (w == 0) ? n0 : (w == 1 ? n1 : n2);
  const float nua = (ua == 0) ? n0 : n1;
  const float nva = (va == 1) ? n1 : n2;
  K.f[F_NW] = nw;
  K.f[F_D1] = S.d1;
  K.f[F_D2] = S.d2;
  K.f[F_NUA] = nua;
  K.f[F_NVA] = nva;
  // plane slab in index space along w: c in [ka*a + kb*b + alpha_lo, ... + alpha_hi]
  if (S.slab) {
    const double inw = 1.0 / (double)nw, iu = 1.0 / (double)P.unit;
    const double a1 = -(S.nc - S.nv0) * inw * iu, a2 = -(S.nm - S.nv0) * inw * iu;
    K.f[F_KA] = (float)(-(double)nua * inw);
    K.f[F_KB] = (float)(-(double)nva * inw);
    K.f[F_ALO] = (float)(fmin(a1, a2) - P.slab_margin);
    K.f[F_AHI] = (float)(fmax(a1, a2) + P.slab_margin);
  } else {
    K.f[F_KA] = K.f[F_KB] = K.f[F_ALO] = K.f[F_AHI] = 0.0f;
  }
  if (w == 0) {
    K.i[I_G0U] = S.g0[1]; K.i[I_G1U] = S.g1[1];
    K.i[I_G0V] = S.g0[2]; K.i[I_G1V] = S.g1[2];
    K.i[I_G0W] = S.g0[0]; K.i[I_G1W] = S.g1[0];
  } else if (w == 1) {
    K.i[I_G0U] = S.g0[0]; K.i[I_G1U] = S.g1[0];
    K.i[I_G0V] = S.g0[2]; K.i[I_G1V] = S.g1[2];
    K.i[I_G0W] = S.g0[1]; K.i[I_G1W] = S.g1[1];
  } else {
    K.i[I_G0U] = S.g0[0]; K.i[I_G1U] = S.g1[0];
    K.i[I_G0V] = S.g0[1]; K.i[I_G1V] = S.g1[1];
    K.i[I_G0W] = S.g0[2]; K.i[I_G1W] = S.g1[2];
  }
  K.i[I_WS] = w | (S.slab ? 4 : 0);
  K.i[I_TRI] = (int)tri;
}

// The voxels of column (a, b) that can pass: false when the column's projection test
// fails, else [lo, hi] along w (the bbox, narrowed to the plane slab). Both passes of
// the voxelizer call this with the same inputs, so they see the same ranges.
__device__ __forceinline__ bool column_range(const Walk& K, float pa, float pb, int a, int b,
                                             int& lo, int& hi) {
#pragma unroll
  for (int k = 0; k < 3; k++)
    if ((K.f[F_H0 + k] * pa + K.f[F_H1 + k] * pb) + K.f[F_HD + k] < 0.0f) return false;
  lo = K.i[I_G0W];
  hi = K.i[I_G1W];
  if (K.slab()) {
    const float base = K.f[F_KA] * (float)a + K.f[F_KB] * (float)b;
    const float l = fminf(fmaxf(base + K.f[F_ALO], (float)lo - 1.0f), (float)hi + 1.0f);
    const float h = fmaxf(fminf(base + K.f[F_AHI], (float)hi + 1.0f), (float)lo - 1.0f);
    lo = max(lo, (int)floorf(l));
    hi = min(hi, (int)floorf(h));
  }
  return lo <= hi;
}

// Pre-pass: flag every brick that a column's candidate voxels fall in (a superset of the
// bricks that will hold marked voxels).
template <typename Act>
__device__ __forceinline__ void touch_column(const Walk& K, const VoxParams& P, int a, int b,
                                             Act& act) {
  const float u = P.unit;
  int lo, hi;
  if (!column_range(K, (float)a * u, (float)b * u, a, b, lo, hi)) return;
  const int w = K.w();
  const int ua = (w == 0) ? 1 : 0;
  const int va = (w == 2) ? 1 : 2;
  const uint64_t bbase = (spread3((uint32_t)(a >> 4)) << ua) | (spread3((uint32_t)(b >> 4)) << va);
  for (int cb = lo >> 4; cb <= (hi >> 4); cb++) act.touch(bbase | (spread3((uint32_t)cb) << w));
}

// Test the voxels of column (a, b) and hand every accepted one to act.add(code, tri).
template <typename Act>
__device__ __forceinline__ void walk_column(const Walk& K, const VoxParams& P, int a, int b,
                                            Act& act) {
  const float u = P.unit;
  const float pa = (float)a * u;
  const float pb = (float)b * u;
  int lo, hi;
  if (!column_range(K, pa, pb, a, b, lo, hi)) return;
  const int w = K.w();
  const float tua = K.f[F_NUA] * pa, tva = K.f[F_NVA] * pb;
  const float Ka = (w == 2) ? (tua + tva) : tua;
  const float Kb = (w == 2) ? 0.0f : tva;
  const float cA = (w == 2) ? pb : pa;
  const float cB = (w == 2) ? pa : pb;
  float KA[3], KB[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    KA[k] = K.f[F_AC + k] * cA;
    KB[k] = K.f[F_BC + k] * cB;
  }
  const int ua = (w == 0) ? 1 : 0;
  const int va = (w == 2) ? 1 : 2;
  const uint64_t cbase = (spread3((uint32_t)a) << ua) | (spread3((uint32_t)b) << va);
  const float nw = K.f[F_NW], d1 = K.f[F_D1], d2 = K.f[F_D2];
  for (int c = lo; c <= hi; c++) {
    const float pc = (float)c * u;
    const float t = (nw * pc + Ka) + Kb;
    if ((t + d1) * (t + d2) > 0.0f) continue;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if ((K.f[F_AV + k] * pc + KA[k]) + K.f[F_AD + k] < 0.0f) ok = false;
      if ((K.f[F_BV + k] * pc + KB[k]) + K.f[F_BD + k] < 0.0f) ok = false;
    }
    if (!ok) continue;
    act.add(cbase | (spread3((uint32_t)c) << w), (uint32_t)K.i[I_TRI]);
  }
}

// Brick allocation without atomics: the touch pass stores 1 into the table entry of
// every brick a column's candidate voxels fall in; a scan over the table then numbers the
// flagged bricks in Morton order (table entry = slot + 1) and lists them.
struct TouchAction {
  static constexpr bool kTouch = true;
  uint32_t* table;
  uint64_t brick0;  // first brick of the partition (morton_start >> kBrickBits)
  __device__ __forceinline__ void touch(uint64_t brick_code) { table[brick_code - brick0] = 1u; }
  __device__ __forceinline__ void add(uint64_t, uint32_t) {}
  __device__ __forceinline__ void flush() {}
};

// Mark voxels: a brick slot holds one byte per voxel (Morton order) and a mark is a plain
// byte store of 1. Concurrent marks of one voxel all store the same value, so no atomic
// read-modify-write is needed however many triangles cover it (about three per voxel on
// a terrain). The brick's slot comes from the table (assigned by the touch pass and the
// table scan), cached per thread.
struct MarkAction {
  static constexpr bool kTouch = false;
  const uint32_t* table;
  uint8_t* pool8;
  ull* ctr;
  uint64_t pool_cap;
  uint64_t mstart;
  uint32_t cbrick = 0xFFFFFFFFu, cslot = 0;
  __device__ __forceinline__ void touch(uint64_t) {}
  __device__ __forceinline__ void add(uint64_t code, uint32_t) {
    const uint64_t rel = code - mstart;
    const uint32_t b = (uint32_t)(rel >> kBrickBits);
    if (b != cbrick) {
      cbrick = b;
      cslot = table[b];
    }
    const uint32_t s = cslot;
    if (s == 0 || s > pool_cap) {
      // s == 0 cannot happen (the touch pass covers every candidate voxel); s > pool_cap
      // means the pool is too small and the pass is redone with a larger one
      atomicAdd(&ctr[s == 0 ? kCtrError : kCtrOverflow], 1ull);
      return;
    }
    pool8[(uint64_t)(s - 1) * kBrickVox + (rel & (kBrickVox - 1))] = 1;
  }
  __device__ __forceinline__ void flush() {}
};

// Second payload pass: lowest triangle index per voxel rank.
struct WinnerAction {
  static constexpr bool kTouch = false;
  __device__ __forceinline__ void touch(uint64_t) {}
  const uint32_t* table;
  const ull* pool;  // bit form of the bricks, written by the compaction
  const uint32_t* wprefix;
  uint32_t* winner;
  ull* ctr;
  uint64_t mstart;
  __device__ __forceinline__ void add(uint64_t code, uint32_t tri) {
    const uint64_t rel = code - mstart;
    const uint32_t b = (uint32_t)(rel >> kBrickBits);
    const uint32_t s = table[b];
    const uint32_t wi = (uint32_t)(rel >> 6) & (kBrickWords - 1);
    const uint64_t slot = (uint64_t)(s - 1) * kBrickWords + wi;
    const ull bit = 1ull << (rel & 63);
    // both passes evaluate the same predicates, so the voxel is marked; a voxel that is
    // not (which would be a bug) is counted and skipped rather than written
    if (s == 0 || !(pool[slot] & bit)) {
      atomicAdd(&ctr[kCtrError], 1ull);
      return;
    }
    const uint32_t rank = wprefix[slot] + (uint32_t)__popcll(pool[slot] & (bit - 1ull));
    atomicMin(&winner[rank], tri);
  }
  __device__ __forceinline__ void flush() {}
};

// Small tier: a block of 256 threads takes 256 consecutive triangles of the partition
// list. Each thread builds its triangle's walk constants into LDS; a block scan lays out
// the (triangle, column) items; then the threads share the items, one column each, so
// the work per lane is a few voxel tests regardless of triangle size. Triangles with more
// than small_cols columns (or non-slab triangles with more than small_cols * 8 tests) go
// to the column-parallel tier.
constexpr int kTile = 256;

template <typename Act>
__global__ __launch_bounds__(kTile) void k_vox_tile(VoxParams P, Act act0, uint32_t* large_tri,
                                                    uint32_t* large_cols, ull* ctr, int collect) {
  __shared__ uint32_t sk[kWalkWords][kTile];
  __shared__ uint32_t soff[kTile + 1];
  __shared__ uint32_t wsum[kTile / kWave];
  Act act = act0;
  const uint32_t tid = threadIdx.x;
  const uint32_t i = blockIdx.x * kTile + tid;
  uint32_t items = 0;
  if (i < P.n) {
    const uint32_t t = P.list ? P.list[i] : i;
    Setup S;
    setup_tri(P.tris + (uint64_t)t * P.stride, P, S);
    Walk K;
    make_walk(S, P, t, K);
    const uint64_t ncols = K.columns();
    const uint64_t lim = (uint64_t)P.small_cols;
    if (K.slab() ? (ncols > lim) : (K.work() > lim * 8)) {
      if (collect) {
        const ull idx = atomicAdd(&ctr[kCtrLargeCount], 1ull);
        large_tri[idx] = t;
        large_cols[idx] = (uint32_t)ncols;
      }
    } else {
      items = (uint32_t)ncols;
#pragma unroll
      for (int f = 0; f < kWalkF; f++) sk[f][tid] = __float_as_uint(K.f[f]);
#pragma unroll
      for (int f = 0; f < kWalkI; f++) sk[kWalkF + f][tid] = (uint32_t)K.i[f];
    }
  }
  uint32_t total;
  const uint32_t inc = block_incl_scan<uint32_t, kTile>(items, wsum, total);
  soff[tid] = inc - items;
  if (tid == kTile - 1) soff[kTile] = total;
  __syncthreads();
  for (uint32_t it = tid; it < total; it += kTile) {
    // owner: last j with soff[j] <= it (triangles without items have empty ranges)
    uint32_t lo = 0, hi = kTile - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (soff[mid] <= it)
        lo = mid;
      else
        hi = mid - 1;
    }
    Walk K;
#pragma unroll
    for (int f = 0; f < kWalkF; f++) K.f[f] = __uint_as_float(sk[f][lo]);
#pragma unroll
    for (int f = 0; f < kWalkI; f++) K.i[f] = (int)sk[kWalkF + f][lo];
    const uint32_t local = it - soff[lo];
    const uint32_t eu = (uint32_t)(K.i[I_G1U] - K.i[I_G0U] + 1);
    const int a = K.i[I_G0U] + (int)(local % eu);
    const int b = K.i[I_G0V] + (int)(local / eu);
    if constexpr (Act::kTouch)
      touch_column(K, P, a, b, act);
    else
      walk_column(K, P, a, b, act);
  }
  act.flush();
}

// Column-parallel tier: one thread per column of the deferred triangles; the counts live
// on the device (ctr[kCtrLargeCount] triangles, ctr[kCtrCols] columns), the grid strides.
template <typename Act>
__global__ __launch_bounds__(256) void k_vox_large(VoxParams P, Act act0,
                                                   const uint32_t* __restrict__ large_tri,
                                                   const uint64_t* __restrict__ large_off,
                                                   const ull* __restrict__ ctr) {
  const uint32_t n_large = (uint32_t)ctr[kCtrLargeCount];
  const uint64_t total = ctr[kCtrCols];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  Act act = act0;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
    // last j with large_off[j] <= g
    uint32_t lo = 0, hi = n_large - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (large_off[mid] <= g)
        lo = mid;
      else
        hi = mid - 1;
    }
    const uint32_t t = large_tri[lo];
    const uint64_t local = g - large_off[lo];
    Setup S;
    setup_tri(P.tris + (uint64_t)t * P.stride, P, S);
    Walk K;
    make_walk(S, P, t, K);
    const uint64_t eu = (uint64_t)(K.i[I_G1U] - K.i[I_G0U] + 1);
    const int a = K.i[I_G0U] + (int)(local % eu);
    const int b = K.i[I_G0V] + (int)(local / eu);
    if constexpr (Act::kTouch)
      touch_column(K, P, a, b, act);
    else
      walk_column(K, P, a, b, act);
  }
  act.flush();
}

// ---------------------------------------------------------------- compaction
struct OccLoad {
  const uint32_t* table;
  __device__ uint32_t operator()(uint64_t i) const { return table[i] != 0u; }
};
struct OccStore {
  uint32_t* occ;
  uint32_t* table;  // flagged entries become slot + 1 (slots in Morton order)
  ull* ctr;
  uint64_t n;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t v) const {
    if (v) {
      occ[ex] = (uint32_t)i;
      table[i] = ex + 1u;
    }
    if (i == n - 1) ctr[kCtrNumOcc] = (ull)(ex + v);
  }
};

// Occupied-brick kernels: one wave per brick (lane = word), striding over the
// ctr[kCtrNumOcc] bricks.
__device__ __forceinline__ uint64_t wave_global() {
  return ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
}
__device__ __forceinline__ uint64_t waves_total() {
  return (uint64_t)gridDim.x * blockDim.x / kWave;
}

// Pack each occupied brick's bytes into 64 bit words (lane l = voxels 64l..64l+63) and
// count its voxels.
__global__ __launch_bounds__(256) void k_brick_count(const uint32_t* __restrict__ occ,
                                                     const ull* __restrict__ ctr,
                                                     uint64_t pool_cap,
                                                     const uint32_t* __restrict__ table,
                                                     const uint8_t* __restrict__ pool8,
                                                     ull* __restrict__ pool,
                                                     uint64_t* __restrict__ bcount) {
  // more flagged bricks than slots: the pass is redone with a larger pool
  const uint64_t n_occ = min((uint64_t)ctr[kCtrNumOcc], pool_cap);
  for (uint64_t wv = wave_global(); wv < n_occ; wv += waves_total()) {
    const uint32_t s = table[occ[wv]];
    const uint4* src = (const uint4*)(pool8 + (uint64_t)(s - 1) * kBrickVox + lane_id() * 64);
    ull m = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint4 v = src[q];
      const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int d = 0; d < 4; d++) {
        const uint32_t x = w4[d];  // four bytes, each 0 or 1
        const uint32_t nib = (x & 1u) | ((x >> 7) & 2u) | ((x >> 14) & 4u) | ((x >> 21) & 8u);
        m |= (ull)nib << (16 * q + 4 * d);
      }
    }
    pool[(uint64_t)(s - 1) * kBrickWords + lane_id()] = m;
    const uint32_t c = (uint32_t)__popcll(m);
    const uint32_t tot = wave_incl_scan<uint32_t>(c);
    if (lane_id() == kWave - 1) bcount[wv] = tot;
  }
}

struct CountStore {
  uint64_t* off;
  ull* total;
  const ull* n_dev;  // actual element count (device)
  __device__ void operator()(uint64_t i, uint64_t ex, uint64_t v) const {
    off[i] = ex;
    if (i == *n_dev - 1) *total = (ull)(ex + v);
  }
};
struct ColLoad {
  const uint32_t* a;
  __device__ uint64_t operator()(uint64_t i) const { return a[i]; }
};

__global__ __launch_bounds__(256) void k_brick_emit(const uint32_t* __restrict__ occ,
                                                    const ull* __restrict__ ctr,
                                                    const uint32_t* __restrict__ table,
                                                    const ull* __restrict__ pool,
                                                    const uint64_t* __restrict__ boff,
                                                    uint64_t mstart, uint64_t* __restrict__ codes,
                                                    uint32_t* __restrict__ wprefix) {
  const uint64_t n_occ = ctr[kCtrNumOcc];
  for (uint64_t wv = wave_global(); wv < n_occ; wv += waves_total()) {
    const uint32_t b = occ[wv];
    const uint32_t s = table[b];
    const uint64_t slot = (uint64_t)(s - 1) * kBrickWords + lane_id();
    ull word = pool[slot];
    const uint32_t c = (uint32_t)__popcll(word);
    const uint32_t inc = wave_incl_scan<uint32_t>(c);
    uint64_t pos = boff[wv] + (inc - c);
    if (wprefix) wprefix[slot] = (uint32_t)pos;
    const uint64_t cb = mstart + ((uint64_t)b << kBrickBits) + ((uint64_t)lane_id() << 6);
    while (word) {
      const int bit = __ffsll(word) - 1;
      codes[pos++] = cb + (uint64_t)bit;
      word &= word - 1;
    }
  }
}

// restore the all-zero state of the table and the byte pool for the occupied bricks
__global__ __launch_bounds__(256) void k_brick_clear(const uint32_t* __restrict__ occ,
                                                     const ull* __restrict__ ctr,
                                                     uint32_t* __restrict__ table,
                                                     uint8_t* __restrict__ pool8) {
  const uint64_t n_occ = ctr[kCtrNumOcc];
  for (uint64_t wv = wave_global(); wv < n_occ; wv += waves_total()) {
    const uint32_t b = occ[wv];
    const uint32_t s = table[b];
    uint4* dst = (uint4*)(pool8 + (uint64_t)(s - 1) * kBrickVox + lane_id() * 64);
#pragma unroll
    for (int q = 0; q < 4; q++) dst[q] = make_uint4(0u, 0u, 0u, 0u);
    if (lane_id() == 0) table[b] = 0u;
  }
}

// ---------------------------------------------------------------- payload records
// BarycentricCoords.h:4-32 for the winning triangle; colour modes of main.cpp:373-382.
__global__ __launch_bounds__(256) void k_payload_records(
    const float* __restrict__ tris, const uint64_t* __restrict__ codes,
    const uint32_t* __restrict__ winner, uint64_t nvox, float unit_div, int color_mode,
    uint64_t gridsize, float* __restrict__ out, ull* ctr) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvox) return;
  const uint64_t code = codes[i];
  const uint32_t t = winner[i];
  float col0, col1, col2, nr0, nr1, nr2;
  if (t == 0xFFFFFFFFu) {  // cannot happen: both passes evaluate the same predicates
    atomicAdd(&ctr[kCtrError], 1ull);
    col0 = col1 = col2 = nr0 = nr1 = nr2 = 0.0f;
  } else {
    const float* p = tris + (uint64_t)t * 21;
    const float v0x = p[0], v0y = p[1], v0z = p[2];
    const float v1x = p[3], v1y = p[4], v1z = p[5];
    const float v2x = p[6], v2y = p[7], v2z = p[8];
    nr0 = p[9];
    nr1 = p[10];
    nr2 = p[11];
    uint32_t x, y, z;
    unmorton3(code, x, y, z);
    if (color_mode == 0) {
      // n = normalize(cross(e0, e1)) exactly as in the voxelizer
      const float e0x = v1x - v0x, e0y = v1y - v0y, e0z = v1z - v0z;
      const float e1x = v2x - v1x, e1y = v2y - v1y, e1z = v2z - v1z;
      const float cx = e0y * e1z - e1y * e0z;
      const float cy = e0z * e1x - e1z * e0x;
      const float cz = e0x * e1y - e1x * e0y;
      const float inv = 1.0f / sqrtf((cx * cx + cy * cy) + cz * cz);
      const float nx = cx * inv, ny = cy * inv, nz = cz * inv;
      const float vx = (float)(int)x / unit_div;
      const float vy = (float)(int)y / unit_div;
      const float vz = (float)(int)z / unit_div;
      const float coeffD = -((v0x * nx + v0y * ny) + v0z * nz);
      const float k = (((vx * nx + vy * ny) + vz * nz) + coeffD) / ((nx * nx + ny * ny) + nz * nz);
      const float qx = vx - k * nx, qy = vy - k * ny, qz = vz - k * nz;
      // glm::inverse(mat3(v0, v1, v2)), m[c][r]
      const float m00 = v0x, m01 = v0y, m02 = v0z;
      const float m10 = v1x, m11 = v1y, m12 = v1z;
      const float m20 = v2x, m21 = v2y, m22 = v2z;
      const float det = (m00 * (m11 * m22 - m21 * m12) - m10 * (m01 * m22 - m21 * m02)) +
                        m20 * (m01 * m12 - m11 * m02);
      const float o = 1.0f / det;
      const float i00 = (m11 * m22 - m21 * m12) * o;
      const float i10 = (-(m10 * m22 - m20 * m12)) * o;
      const float i20 = (m10 * m21 - m20 * m11) * o;
      const float i01 = (-(m01 * m22 - m21 * m02)) * o;
      const float i11 = (m00 * m22 - m20 * m02) * o;
      const float i21 = (-(m00 * m21 - m20 * m01)) * o;
      const float i02 = (m01 * m12 - m11 * m02) * o;
      const float i12 = (-(m00 * m12 - m10 * m02)) * o;
      const float i22 = (m00 * m11 - m10 * m01) * o;
      const float bx = (i00 * qx + i10 * qy) + i20 * qz;
      const float by = (i01 * qx + i11 * qy) + i21 * qz;
      const float bz = (i02 * qx + i12 * qy) + i22 * qz;
      col0 = (bx * p[12] + by * p[15]) + bz * p[18];
      col1 = (bx * p[13] + by * p[16]) + bz * p[19];
      col2 = (bx * p[14] + by * p[17]) + bz * p[20];
    } else if (color_mode == 1) {  // fixed white
      col0 = col1 = col2 = 1.0f;
    } else if (color_mode == 2) {  // mortonToRGB: (z, y, x) / gridsize
      const float g = (float)gridsize;
      col0 = (float)z / g;
      col1 = (float)y / g;
      col2 = (float)x / g;
    } else {  // normal -> colour
      const float inv = 1.0f / sqrtf((nr0 * nr0 + nr1 * nr1) + nr2 * nr2);
      const float a = nr0 * inv, b = nr1 * inv, c = nr2 * inv;
      col0 = (a + 1.0f) / 2.0f;
      col1 = (b + 1.0f) / 2.0f;
      col2 = (c + 1.0f) / 2.0f;
    }
  }
  float* r = out + i * 8;
  *(uint64_t*)r = code;
  r[2] = canon_nan(col0);
  r[3] = canon_nan(col1);
  r[4] = canon_nan(col2);
  r[5] = canon_nan(nr0);
  r[6] = canon_nan(nr1);
  r[7] = canon_nan(nr2);
}

// ---------------------------------------------------------------- partition kernels
// Per triangle, the box [lo, hi] of partition cube coordinates its bbox touches, packed
// as 8-bit fields (x lo, x hi, y lo, y hi, z lo, z hi); ~0 = touches none.
__global__ __launch_bounds__(256) void k_part_hitbox(const float* __restrict__ tris, int stride,
                                                     uint64_t n,
                                                     const float* __restrict__ cube_lo,
                                                     const float* __restrict__ cube_hi, int K,
                                                     uint64_t* __restrict__ hitbox) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float* p = tris + t * stride;
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
  hitbox[t] = none ? ~0ull : packed;
}

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
struct PartStore {
  uint32_t* list;
  ull* total;
  uint64_t n;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t v) const {
    if (v) list[ex] = (uint32_t)i;
    if (i == n - 1) *total = (ull)(ex + v);
  }
};

// ---------------------------------------------------------------- fills
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}
static unsigned fill_grid(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 8192));
}
hipError_t fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}
hipError_t fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------- host drivers
#define SVO_TRY(x)                        \
  do {                                    \
    hipError_t _e = (x);                  \
    if (_e != hipSuccess) return _e;      \
  } while (0)

static hipError_t read_counters(Context& c, ull* h) {
  SVO_TRY(hipMemcpyAsync(h, c.counters.p, kNumCounters * sizeof(ull), hipMemcpyDeviceToHost,
                         c.stream));
  return hipStreamSynchronize(c.stream);
}

static hipError_t ensure_scan_scratch(Context& c, uint64_t n, size_t elem) {
  return ensure(c.scan_scratch, scan_scratch_elems(n) * elem + 256);
}

hipError_t partition_mesh(Context& c, float unitlength, uint64_t gridsize, uint64_t n_parts) {
  c.n_parts = n_parts;
  c.part_off.assign(n_parts, 0);
  c.part_cnt.assign(n_parts, 0);
  if (n_parts <= 1) {
    c.part_cnt[0] = c.n_tris;
    return hipSuccess;
  }
  // partitions are the 8^k aligned Morton cubes; K cubes per axis
  int k = 0;
  while ((1ull << (3 * k)) < n_parts) k++;
  const int K = 1 << k;
  if (K > 256) {
    set_error("more than 256 partitions per axis are not supported");
    return hipErrorInvalidValue;
  }
  const uint64_t side = gridsize / (uint64_t)K;
  // world box of cube c along any axis: [c*side*u, (c+1)*side*u] (partitioner.cpp:56-61)
  std::vector<float> lo(K), hi(K);
  for (int i = 0; i < K; i++) {
    lo[i] = (float)(uint32_t)(i * side) * unitlength;
    hi[i] = (float)(uint32_t)((i + 1) * side) * unitlength;
  }
  const int stride = c.payload ? 21 : 9;
  SVO_TRY(ensure(c.hitbox, c.n_tris * 8 + 2 * 256 * sizeof(float) + 256));
  float* d_lo = (float*)((char*)c.hitbox.p + c.n_tris * 8 + 64);
  float* d_hi = d_lo + 256;
  SVO_TRY(hipMemcpyAsync(d_lo, lo.data(), K * 4, hipMemcpyHostToDevice, c.stream));
  SVO_TRY(hipMemcpyAsync(d_hi, hi.data(), K * 4, hipMemcpyHostToDevice, c.stream));
  if (c.n_tris) {
    ktimer_begin(c, "part_hitbox");
    hipLaunchKernelGGL(k_part_hitbox, dim3((unsigned)((c.n_tris + 255) / 256)), dim3(256), 0,
                       c.stream, c.tris.as<float>(), stride, c.n_tris, d_lo, d_hi, K,
                       c.hitbox.as<uint64_t>());
    ktimer_end(c);
    SVO_TRY(hipGetLastError());
  }
  // one order-preserving flag scan per partition; lists concatenated in partition order
  SVO_TRY(ensure_scan_scratch(c, c.n_tris, 4));
  std::vector<ull> totals(n_parts, 0);
  SVO_TRY(ensure(c.part_lists, std::max<uint64_t>(c.n_tris, 1) * 4 * 2));
  uint64_t used = 0;
  ull* d_tot = c.counters.as<ull>() + 16;
  for (uint64_t j = 0; j < n_parts; j++) {
    // list capacity: grow if the previous partitions filled it
    if ((used + c.n_tris) * 4 > c.part_lists.cap)
      SVO_TRY(ensure_keep(c.part_lists, (used + c.n_tris) * 4 * 2, used * 4, c.stream));
    SVO_TRY(fill_u64((uint64_t*)d_tot, 0, 1, c.stream));
    uint64_t cx = 0, cy = 0, cz = 0;
    {
      uint32_t x, y, z;
      unmorton3(j, x, y, z);
      cx = x;
      cy = y;
      cz = z;
    }
    if (c.n_tris) {
      PartLoad ld{c.hitbox.as<uint64_t>(), (uint32_t)cx, (uint32_t)cy, (uint32_t)cz};
      PartStore st{c.part_lists.as<uint32_t>() + used, d_tot, c.n_tris};
      SVO_TRY((scan_exclusive<uint32_t>(ld, st, c.n_tris, c.scan_scratch.as<uint32_t>(), c.stream)));
    }
    SVO_TRY(hipMemcpyAsync(&totals[j], d_tot, 8, hipMemcpyDeviceToHost, c.stream));
    SVO_TRY(hipStreamSynchronize(c.stream));
    c.part_off[j] = used;
    c.part_cnt[j] = totals[j];
    used += totals[j];
  }
  return hipSuccess;
}

hipError_t voxelize_partition(Context& c, uint64_t part, uint64_t morton_start,
                              uint64_t morton_end, float unitlength, uint64_t gridsize,
                              int color_mode, uint64_t* nfilled) {
  const uint64_t n = (c.n_parts > 1) ? c.part_cnt[part] : c.n_tris;
  *nfilled = 0;
  if (n == 0) return hipSuccess;
  if (n > 0xFFFFFFF0ull) {
    set_error("partition holds more than 2^32 triangles");
    return hipErrorInvalidValue;
  }
  VoxParams P;
  P.tris = c.tris.as<float>();
  P.list = (c.n_parts > 1) ? c.part_lists.as<uint32_t>() + c.part_off[part] : nullptr;
  P.n = (uint32_t)n;
  P.stride = c.payload ? 21 : 9;
  P.mstart = morton_start;
  {
    uint32_t x, y, z;
    unmorton3(morton_start, x, y, z);
    P.pmin[0] = (int)x;
    P.pmin[1] = (int)y;
    P.pmin[2] = (int)z;
    unmorton3(morton_end - 1, x, y, z);
    P.pmax[0] = (int)x;
    P.pmax[1] = (int)y;
    P.pmax[2] = (int)z;
  }
  P.unit = unitlength;
  P.unit_div = 1.0f / unitlength;  // voxelizer.cpp:164, IEEE on the host as on the reference
  P.coord_limit = 4.0f * (float)gridsize * unitlength;
  // The reference's FP32 plane test (t + d1) * (t + d2) differs from its exact value by
  // less than ~3e-6 * coord_limit; along the column axis (|n_w| >= 1/2 here) that is
  // below 2.1e-5 * G voxels. The margin is 0.5 + 4e-5 * G voxels, at least twice that.
  P.slab_margin = 0.5 + 4e-5 * (double)gridsize;
  P.small_cols = c.tune.small_cols;

  const uint64_t span = morton_end - morton_start;
  const uint64_t nbricks = (span + (1ull << kBrickBits) - 1) >> kBrickBits;
  if (nbricks > 0xFFFFFFF0ull) {
    set_error("partition too large for the brick table");
    return hipErrorInvalidValue;
  }
  // the table and the pool are all-zero whenever no pass is running
  if (c.table_entries < nbricks) {
    SVO_TRY(ensure(c.table, nbricks * 4));
    SVO_TRY(fill_u32(c.table.as<uint32_t>(), 0, nbricks, c.stream));
    c.table_entries = nbricks;
  }
  if (c.pool_cap == 0) {
    // lost install races leave unused slots behind, so the pool is not capped at nbricks
    const uint64_t cap = std::max<uint64_t>(1024, n / 64);
    SVO_TRY(ensure(c.pool, cap * kBrickWords * 8));
    SVO_TRY(ensure(c.pool8, cap * kBrickVox));
    SVO_TRY(fill_u64(c.pool8.as<uint64_t>(), 0, cap * kBrickVox / 8, c.stream));
    c.pool_cap = cap;
  }
  SVO_TRY(ensure(c.large_tri, n * 4));
  SVO_TRY(ensure(c.large_cols, n * 4));
  SVO_TRY(ensure(c.large_off, n * 8 + 8));
  SVO_TRY(ensure(c.occ_list, nbricks * 4));
  SVO_TRY(ensure_scan_scratch(c, std::max<uint64_t>(n, nbricks), 8));
  ull* ctr = c.counters.as<ull>();
  ull hc[kNumCounters];
  const unsigned nblk = (unsigned)((n + 255) / 256);
  const unsigned large_grid = 2048;
  auto brick_grid = [&](uint64_t bricks) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((bricks * kWave + 255) / 256, 8192));
  };

  // ---- phase A: mark, compact counts; one synchronization at its end
  for (int attempt = 0;; attempt++) {
    const uint64_t bc_bound = std::min<uint64_t>(nbricks, c.pool_cap);
    SVO_TRY(ensure(c.bcount, bc_bound * 8 + 8));
    SVO_TRY(fill_u64((uint64_t*)ctr, 0, 7, c.stream));  // counters 0..6
    // touch pass: flag the bricks of every candidate voxel, collect the large triangles
    TouchAction touch{c.table.as<uint32_t>(), morton_start >> kBrickBits};
    ktimer_begin(c, "vox_small_touch");
    hipLaunchKernelGGL((k_vox_tile<TouchAction>), dim3(nblk), dim3(kTile), 0, c.stream, P, touch,
                       c.large_tri.as<uint32_t>(), c.large_cols.as<uint32_t>(), ctr, 1);
    ktimer_end(c);
    SVO_TRY(hipGetLastError());
    SVO_TRY((scan_exclusive<uint64_t>(
        ColLoad{c.large_cols.as<uint32_t>()},
        CountStore{c.large_off.as<uint64_t>(), ctr + kCtrCols, ctr + kCtrLargeCount}, n,
        c.scan_scratch.as<uint64_t>(), c.stream, ctr + kCtrLargeCount)));
    ktimer_begin(c, "vox_large_touch");
    hipLaunchKernelGGL((k_vox_large<TouchAction>), dim3(large_grid), dim3(256), 0, c.stream, P,
                       touch, c.large_tri.as<uint32_t>(), c.large_off.as<uint64_t>(), ctr);
    ktimer_end(c);
    SVO_TRY(hipGetLastError());
    // number the flagged bricks in Morton order: table entry = slot + 1, occ = brick list
    SVO_TRY((scan_exclusive<uint32_t>(
        OccLoad{c.table.as<uint32_t>()},
        OccStore{c.occ_list.as<uint32_t>(), c.table.as<uint32_t>(), ctr, nbricks}, nbricks,
        c.scan_scratch.as<uint32_t>(), c.stream)));
    // mark pass: byte stores into the slots
    MarkAction mark{c.table.as<uint32_t>(), c.pool8.as<uint8_t>(), ctr, c.pool_cap, morton_start};
    ktimer_begin(c, "vox_small_mark");
    hipLaunchKernelGGL((k_vox_tile<MarkAction>), dim3(nblk), dim3(kTile), 0, c.stream, P, mark,
                       c.large_tri.as<uint32_t>(), c.large_cols.as<uint32_t>(), ctr, 0);
    ktimer_end(c);
    SVO_TRY(hipGetLastError());
    ktimer_begin(c, "vox_large_mark");
    hipLaunchKernelGGL((k_vox_large<MarkAction>), dim3(large_grid), dim3(256), 0, c.stream, P,
                       mark, c.large_tri.as<uint32_t>(), c.large_off.as<uint64_t>(), ctr);
    ktimer_end(c);
    SVO_TRY(hipGetLastError());
    // per-brick voxel counts (bytes packed to bits) and their offsets
    ktimer_begin(c, "brick_count");
    hipLaunchKernelGGL(k_brick_count, dim3(brick_grid(bc_bound)), dim3(256), 0, c.stream,
                       c.occ_list.as<uint32_t>(), ctr, c.pool_cap, c.table.as<uint32_t>(),
                       c.pool8.as<uint8_t>(), c.pool.as<ull>(), c.bcount.as<uint64_t>());
    ktimer_end(c);
    SVO_TRY(hipGetLastError());
    SVO_TRY((scan_exclusive<uint64_t>(
        ArrayLoad<uint64_t>{c.bcount.as<uint64_t>()},
        CountStore{c.bcount.as<uint64_t>(), ctr + kCtrNumVox, ctr + kCtrNumOcc}, bc_bound,
        c.scan_scratch.as<uint64_t>(), c.stream, ctr + kCtrNumOcc)));
    SVO_TRY(read_counters(c, hc));
    ktimer_flush(c);
    if (!hc[kCtrOverflow] && hc[kCtrNumOcc] <= c.pool_cap) break;
    // pool too small for the flagged bricks: clear everything, grow, redo
    const uint64_t want = std::max<uint64_t>(c.pool_cap * 2, hc[kCtrNumOcc] + 1024);
    SVO_TRY(fill_u32(c.table.as<uint32_t>(), 0, nbricks, c.stream));
    release(c.pool);
    release(c.pool8);
    SVO_TRY(ensure(c.pool, want * kBrickWords * 8));
    SVO_TRY(ensure(c.pool8, want * kBrickVox));
    SVO_TRY(fill_u64(c.pool8.as<uint64_t>(), 0, want * kBrickVox / 8, c.stream));
    c.pool_cap = want;
    if (attempt > 8 || want > (1ull << 26)) {
      set_error("brick pool could not be sized");
      return hipErrorOutOfMemory;
    }
  }
  const uint64_t n_occ = hc[kCtrNumOcc];
  const uint64_t nvox = hc[kCtrNumVox];
  const uint64_t n_large = hc[kCtrLargeCount];

  // ---- phase B: emit sorted codes, payload, clear; no synchronization
  if (n_occ) {
    if (c.payload) SVO_TRY(ensure(c.wprefix, c.pool_cap * kBrickWords * 4));
    SVO_TRY(ensure_keep(c.codes, (c.n_vox + nvox) * 8, c.n_vox * 8, c.stream));
    ktimer_begin(c, "brick_emit");
    hipLaunchKernelGGL(k_brick_emit, dim3(brick_grid(n_occ)), dim3(256), 0, c.stream,
                       c.occ_list.as<uint32_t>(), ctr, c.table.as<uint32_t>(), c.pool.as<ull>(),
                       c.bcount.as<uint64_t>(), morton_start, c.codes.as<uint64_t>() + c.n_vox,
                       c.payload ? c.wprefix.as<uint32_t>() : (uint32_t*)nullptr);
    ktimer_end(c);
    SVO_TRY(hipGetLastError());

    if (c.payload) {
      if (nvox > 0xFFFFFFF0ull) {
        set_error("partition holds more than 2^32 voxels");
        return hipErrorInvalidValue;
      }
      SVO_TRY(ensure(c.winner, nvox * 4));
      SVO_TRY(fill_u32(c.winner.as<uint32_t>(), 0xFFFFFFFFu, nvox, c.stream));
      WinnerAction win{c.table.as<uint32_t>(), c.pool.as<ull>(), c.wprefix.as<uint32_t>(),
                       c.winner.as<uint32_t>(), ctr, morton_start};
      ktimer_begin(c, "vox_small_winner");
      hipLaunchKernelGGL((k_vox_tile<WinnerAction>), dim3(nblk), dim3(kTile), 0, c.stream, P,
                         win, c.large_tri.as<uint32_t>(), c.large_cols.as<uint32_t>(), ctr, 0);
      ktimer_end(c);
      SVO_TRY(hipGetLastError());
      if (n_large) {
        ktimer_begin(c, "vox_large_winner");
        hipLaunchKernelGGL((k_vox_large<WinnerAction>), dim3(large_grid), dim3(256), 0,
                           c.stream, P, win, c.large_tri.as<uint32_t>(),
                           c.large_off.as<uint64_t>(), ctr);
        ktimer_end(c);
        SVO_TRY(hipGetLastError());
      }
      // records: slot 0 of c.records is the null record
      SVO_TRY(ensure_keep(c.records, (1 + c.n_vox + nvox) * 32, (1 + c.n_vox) * 32, c.stream));
      ktimer_begin(c, "payload_records");
      hipLaunchKernelGGL(k_payload_records, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0,
                         c.stream, c.tris.as<float>(), c.codes.as<uint64_t>() + c.n_vox,
                         c.winner.as<uint32_t>(), nvox, P.unit_div, color_mode, gridsize,
                         (float*)((char*)c.records.p + (1 + c.n_vox) * 32), ctr);
      ktimer_end(c);
      SVO_TRY(hipGetLastError());
    }
    ktimer_begin(c, "brick_clear");
    hipLaunchKernelGGL(k_brick_clear, dim3(brick_grid(n_occ)), dim3(256), 0, c.stream,
                       c.occ_list.as<uint32_t>(), ctr, c.table.as<uint32_t>(),
                       c.pool8.as<uint8_t>());
    ktimer_end(c);
    SVO_TRY(hipGetLastError());
  }
  c.n_vox += nvox;
  *nfilled = nvox;
  return hipSuccess;
}

}  // namespace svo
