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
// on the position along the column), and only the voxels of the plane slab, widened by
// a margin, are tested; every voxel outside that range fails the reference's plane
// test (the margin is over four times the FP32 error of the test). Triangles with a
// non-finite normal or far-away vertices walk full columns.
// Marked voxels go to a sparse brick grid; compaction of the occupied bricks in Morton
// order yields the sorted unique codes the builder needs (the reference sorts its side
// array, main.cpp:356-359). Payload voxels are resolved to the first triangle in
// partition order that covers them (voxelizer.cpp:263) by a second pass doing
// atomicMin on the triangle index per voxel rank.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>

#include "svo_device.h"
#include "svo_internal.h"
#include "svo_scan.h"

namespace svo {

typedef unsigned long long ull;

// One voxelize unit: one Morton-range partition of one job (a mesh). A batch voxelizes
// many units in the same launches.
struct VoxParams {
  const float* tris;
  const uint32_t* list;  // partition's triangle indices, nullptr = identity
  uint32_t n;
  int stride;  // floats per record: 9 binary, 21 payload
  uint64_t mstart;
  uint64_t bstart;  // mstart rounded down to a brick: origin of the brick table segment
  int pmin[3], pmax[3];
  float unit, unit_div;
  float coord_limit;  // |vertex| beyond this walks full columns
  double slab_margin; // voxels added on each side of the exact plane slab
  double inv_unit;    // 1 / unit in double
  int small_cols;
  int direct_cols;     // triangles with at most this many columns: walked by their thread
  uint32_t job;        // tag of the unit's codes (bits kJobShift and up)
  int dense;           // 1: voxels in grid8; 0: in pool slots flagged by the touch pass
  int wide;            // dense payload unit: 4-byte cells holding ~(first triangle)
  uint64_t table_base; // first entry of the unit's segment of the brick table
  uint8_t* grid8;      // dense byte grid of the unit (nullptr on the touch path)
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
//   * each edge of the two other planes is (cv*pc + cc*pconst) + de per voxel, where
//     pconst is pb for one of them (plane P) and pa for the other (plane Q).
// Every product and sum is the reference's own (voxelizer.cpp:257-287); only operands of
// single additions are swapped, which is exact in IEEE arithmetic. The products of pb
// (h1*pb, cc*pb of plane P, the plane-test term of b, the slab term of b) are the same
// along a row of columns and are computed once per row.
enum WalkF {
  F_NW = 0, F_D1, F_D2, F_NUA, F_NVA,
  F_H0 = 5, F_H1 = 8, F_HD = 11,
  F_PV = 14, F_PC = 17, F_PD = 20,
  F_QV = 23, F_QC = 26, F_QD = 29,
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

// The plane roles for column axis w, as value selects: the hoisted plane's edges (H0, H1
// in operand order (pa, pb), HD), plane P (its other coordinate is pb) and plane Q (pa)
// with their column-coordinate factor (PV, QV), the other factor (PC, QC) and the offset
// (PD, QD). Edges: xy = ne[0..2], yz = ne[3..5], zx = ne[6..8].
//   w = 2: hoisted xy(pa, pb); P = yz(pb, pc); Q = zx(pc, pa)
//   w = 0: hoisted yz(pa, pb); P = zx(pb, pc); Q = xy(pc, pa)
//   w = 1: hoisted zx(pb, pa); P = yz(pc, pb); Q = xy(pa, pc)
// (A branch per axis is merged by the compiler into an index on the edge arrays, which
// puts them in scratch memory.)
__device__ __forceinline__ void fill_planes(const Setup& S, Walk& K, int w) {
  const bool w2 = (w == 2), w0 = (w == 0);
  auto pick = [&](float a2, float a0, float a1) { return w2 ? a2 : (w0 ? a0 : a1); };
#pragma unroll
  for (int k = 0; k < 3; k++) {
    K.f[F_H0 + k] = pick(S.ne[k][0], S.ne[3 + k][0], S.ne[6 + k][1]);
    K.f[F_H1 + k] = pick(S.ne[k][1], S.ne[3 + k][1], S.ne[6 + k][0]);
    K.f[F_HD + k] = pick(S.de[k], S.de[3 + k], S.de[6 + k]);
    K.f[F_PV + k] = pick(S.ne[3 + k][1], S.ne[6 + k][1], S.ne[3 + k][0]);
    K.f[F_PC + k] = pick(S.ne[3 + k][0], S.ne[6 + k][0], S.ne[3 + k][1]);
    K.f[F_PD + k] = pick(S.de[3 + k], S.de[6 + k], S.de[3 + k]);
    K.f[F_QV + k] = pick(S.ne[6 + k][0], S.ne[k][0], S.ne[k][1]);
    K.f[F_QC + k] = pick(S.ne[6 + k][1], S.ne[k][1], S.ne[k][0]);
    K.f[F_QD + k] = pick(S.de[6 + k], S.de[k], S.de[k]);
  }
}

// Re-slot a triangle's setup (voxelizer.cpp:189-254, see setup_tri) for the column walk.
__device__ __forceinline__ void make_walk(const Setup& S, const VoxParams& P, uint32_t tri,
                                          Walk& K) {
  const int w = S.w;
  const int ua = (w == 0) ? 1 : 0;
  const int va = (w == 2) ? 1 : 2;
  fill_planes(S, K, w);
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
    const double inw = 1.0 / (double)nw, iu = P.inv_unit;
    const double a1 = -(S.nc - S.nv0) * inw * iu, a2 = -(S.nm - S.nv0) * inw * iu;
    K.f[F_KA] = (float)(-(double)nua * inw);
    K.f[F_KB] = (float)(-(double)nva * inw);
    K.f[F_ALO] = (float)(fmin(a1, a2) - P.slab_margin);
    K.f[F_AHI] = (float)(fmax(a1, a2) + P.slab_margin);
  } else {
    K.f[F_KA] = K.f[F_KB] = K.f[F_ALO] = K.f[F_AHI] = 0.0f;
  }
  // bbox along (ua, va, w) by value selects (a branch per axis would let the compiler
  // index the bbox arrays with w, which puts them in scratch memory)
  const int x0 = S.g0[0], y0 = S.g0[1], z0 = S.g0[2];
  const int x1 = S.g1[0], y1 = S.g1[1], z1 = S.g1[2];
  K.i[I_G0U] = (w == 0) ? y0 : x0;
  K.i[I_G1U] = (w == 0) ? y1 : x1;
  K.i[I_G0V] = (w == 2) ? y0 : z0;
  K.i[I_G1V] = (w == 2) ? y1 : z1;
  K.i[I_G0W] = (w == 0) ? x0 : (w == 1 ? y0 : z0);
  K.i[I_G1W] = (w == 0) ? x1 : (w == 1 ? y1 : z1);
  K.i[I_WS] = w | (S.slab ? 4 : 0);
  K.i[I_TRI] = (int)tri;
}

// The terms of a row of columns (fixed b) that do not depend on a: pb, the hoisted
// plane's h1*pb, plane P's cc*pb, the plane test's term of pb and the slab's term of b.
struct ColRow {
  float pb;
  float hb[3];
  float kp[3];
  float tvb;
  float sb;
};
__device__ __forceinline__ ColRow col_row(const Walk& K, const VoxParams& P, int b) {
  ColRow r;
  r.pb = (float)b * P.unit;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    r.hb[k] = K.f[F_H1 + k] * r.pb;
    r.kp[k] = K.f[F_PC + k] * r.pb;
  }
  r.tvb = K.f[F_NVA] * r.pb;
  r.sb = K.f[F_KB] * (float)b;
  return r;
}

// Rows at least this many columns wide get their columns narrowed to the interval that
// can pass the projection test (row_interval); narrower rows walk every column.
constexpr int kRowSpanMin = 5;

// Columns [a0, a1] of row b: a superset of the columns whose hoisted projection test
// passes (every column outside fails it). An edge's test value at column a,
// fl(fl(fl(h0 * fl(a * u)) + hb) + hd), differs from the real h0 u a + hb + hd by at most
// 5 eps (|h0| |a u| + |hb| + |hd|) (four roundings, eps = 2^-24); with E above that bound,
// a column that passes has h0 u a >= -(hb + hd) - E, a half-line in a. Its end is
// widened by one column for the rounding of its own evaluation (an approximate
// reciprocal, well under a column at these magnitudes). Non-finite constants narrow
// nothing.
__device__ __forceinline__ void row_interval(const Walk& K, const ColRow& R, float u, int& a0,
                                             int& a1) {
  a0 = K.i[I_G0U];
  a1 = K.i[I_G1U];
  if (a1 - a0 + 1 < kRowSpanMin) return;
  const float flo = (float)a0 - 1.0f, fhi = (float)a1 + 1.0f;
  const float amax = fmaxf(fabsf(flo), fabsf(fhi));
  float lo = flo, hi = fhi;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float h0 = K.f[F_H0 + k], hb = R.hb[k], hd = K.f[F_HD + k];
    const float E = 1e-6f * ((fabsf(h0) * amax) * fabsf(u) + fabsf(hb) + fabsf(hd)) + 1e-30f;
    const float rhs = -(hb + hd) - E;
    const float hu = h0 * u;
    if (!(fabsf(rhs) <= 3.0e38f) || !(fabsf(hu) <= 3.0e38f)) continue;
    if (hu > 0.0f)
      lo = fmaxf(lo, rhs * __builtin_amdgcn_rcpf(hu) - 1.0f);
    else if (hu < 0.0f)
      hi = fminf(hi, rhs * __builtin_amdgcn_rcpf(hu) + 1.0f);
    else if (rhs > 0.0f)  // h0 u = 0: no column of the row passes
      hi = flo;
  }
  a0 = max(a0, (int)ceilf(__builtin_amdgcn_fmed3f(lo, flo, fhi)));
  a1 = min(a1, (int)floorf(__builtin_amdgcn_fmed3f(hi, flo, fhi)));
}

// The voxels of column (a, b) that can pass: false when the column's projection test
// fails, else [lo, hi] along w (the bbox, narrowed to the plane slab). Both passes of
// the voxelizer call this with the same inputs, so they see the same ranges.
__device__ __forceinline__ bool column_range(const Walk& K, float pa, const ColRow& R, int a,
                                             int& lo, int& hi) {
  // the three edge tests evaluated together, one branch (their outcome, not their order,
  // decides)
  bool in = true;
#pragma unroll
  for (int k = 0; k < 3; k++)
    in &= !((K.f[F_H0 + k] * pa + R.hb[k]) + K.f[F_HD + k] < 0.0f);
  if (!in) return false;
  lo = K.i[I_G0W];
  hi = K.i[I_G1W];
  if (K.slab()) {
    // the slab ends, clamped to [lo - 1, hi + 1] (finite here: the slab is only used
    // for finite normals and vertices) before the conversion to int
    const float base = K.f[F_KA] * (float)a + R.sb;
    const float l = __builtin_amdgcn_fmed3f(base + K.f[F_ALO], (float)lo - 1.0f, (float)hi + 1.0f);
    const float h = __builtin_amdgcn_fmed3f(base + K.f[F_AHI], (float)lo - 1.0f, (float)hi + 1.0f);
    lo = max(lo, (int)ceilf(l));
    hi = min(hi, (int)floorf(h));
  }
  return lo <= hi;
}

// Pre-pass: flag every brick that a column's candidate voxels fall in (a superset of the
// bricks that will hold marked voxels).
template <typename Act>
__device__ __forceinline__ void touch_column(const Walk& K, const VoxParams& P, int a,
                                             const ColRow& R, int b, Act& act) {
  int lo, hi;
  if (!column_range(K, (float)a * P.unit, R, a, lo, hi)) return;
  const int w = K.w();
  const int ua = (w == 0) ? 1 : 0;
  const int va = (w == 2) ? 1 : 2;
  const uint64_t bbase = (spread3((uint32_t)(a >> 4)) << ua) | (spread3((uint32_t)(b >> 4)) << va);
  for (int cb = lo >> 4; cb <= (hi >> 4); cb++) act.touch(bbase | (spread3((uint32_t)cb) << w));
}

// Morton spreading in the walks from a table in LDS (s_spread[v] = spread3(v) for 8-bit
// v): three lookups and a few shifts instead of the five-step bit spread. Every kernel
// that walks columns fills the table with fill_spread_table before its first walk.
__shared__ uint32_t s_spread[256];
__device__ __forceinline__ void fill_spread_table() {
  for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x) s_spread[i] = (uint32_t)spread3(i);
}
// spread3(v) for v < 2^21
__device__ __forceinline__ uint64_t spread3_s(uint32_t v) {
  const uint32_t a = s_spread[v & 255u], b = s_spread[(v >> 8) & 255u], c = s_spread[(v >> 16) & 31u];
  return (uint64_t)(a | (b << 24)) | ((uint64_t)((b >> 8) | (c << 16)) << 32);
}

// Every bit of Morton axis k: a coordinate spread onto these bits advances by one with a
// masked add, (m | ~mask) + (1 << k), then & mask.
constexpr uint64_t kAxis0Bits = 0x1249249249249249ull;
__device__ __forceinline__ uint64_t morton_step(uint64_t m, int k) {
  const uint64_t mask = kAxis0Bits << k;
  return ((m | ~mask) + (1ull << k)) & mask;
}

// Test the voxels of column (a, b) and hand every accepted one to act.add(code, tri);
// cbase holds a and b spread onto their Morton axes.
template <typename Act>
__device__ __forceinline__ void walk_column(const Walk& K, const VoxParams& P, int a,
                                            const ColRow& R, uint64_t cbase, Act& act) {
  const float u = P.unit;
  const float pa = (float)a * u;
  int lo, hi;
  if (!column_range(K, pa, R, a, lo, hi)) return;
  const int w = K.w();
  const float tua = K.f[F_NUA] * pa;
  const float Ka = (w == 2) ? (tua + R.tvb) : tua;
  const float Kb = (w == 2) ? 0.0f : R.tvb;
  float KQ[3];
#pragma unroll
  for (int k = 0; k < 3; k++) KQ[k] = K.f[F_QC + k] * pa;
  const float nw = K.f[F_NW], d1 = K.f[F_D1], d2 = K.f[F_D2];
  uint64_t mc = spread3_s((uint32_t)lo) << w;
  for (int c = lo; c <= hi; c++, mc = morton_step(mc, w)) {
    const float pc = (float)c * u;
    const float t = (nw * pc + Ka) + Kb;
    // all seven tests, then one branch
    bool ok = !((t + d1) * (t + d2) > 0.0f);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      ok &= !((K.f[F_PV + k] * pc + R.kp[k]) + K.f[F_PD + k] < 0.0f);
      ok &= !((K.f[F_QV + k] * pc + KQ[k]) + K.f[F_QD + k] < 0.0f);
    }
    if (!ok) continue;
    // cbase and mc occupy disjoint bits: the sum is their union, and an action's
    // loop-invariant offset from cbase folds out of the loop
    act.add(cbase + mc, (uint32_t)K.i[I_TRI]);
  }
}

// Columns [i0, i1] of row b (ma = i0 spread onto axis ua, mb = b onto axis va).
template <typename Act>
__device__ __forceinline__ void walk_row_cols(const Walk& K, const VoxParams& P, int b,
                                              const ColRow& R, int i0, int i1, uint64_t ma,
                                              uint64_t mb, int ua, Act& act) {
  for (int a = i0; a <= i1; a++, ma = morton_step(ma, ua)) {
    if constexpr (Act::kTouch)
      touch_column(K, P, a, R, b, act);
    else
      walk_column(K, P, a, R, ma | mb, act);
  }
}

// Columns [a_lo, a_hi] of row b of a walk, narrowed to the row's interval when kNarrow
// (a span of one or a few columns is cheaper to test column by column).
template <bool kNarrow, typename Act>
__device__ __forceinline__ void walk_row_span(const Walk& K, const VoxParams& P, int b, int a_lo,
                                              int a_hi, Act& act) {
  const ColRow R = col_row(K, P, b);
  int i0 = a_lo, i1 = a_hi;
  if constexpr (kNarrow) {
    row_interval(K, R, P.unit, i0, i1);
    i0 = max(i0, a_lo);
    i1 = min(i1, a_hi);
    if (i0 > i1) return;
  }
  const int w = K.w();
  const int ua = (w == 0) ? 1 : 0;
  const int va = (w == 2) ? 1 : 2;
  walk_row_cols(K, P, b, R, i0, i1, spread3_s((uint32_t)i0) << ua,
                spread3_s((uint32_t)b) << va, ua, act);
}

// All columns of a walk, row by row, the Morton parts of a and b stepped incrementally;
// with 'narrow' (units whose items are segments of several columns) rows are narrowed to
// their interval (rows narrower than kRowSpanMin columns are walked whole).
template <typename Act>
__device__ __forceinline__ void walk_all_columns(const Walk& K, const VoxParams& P, Act& act,
                                                 bool narrow) {
  const int w = K.w();
  const int ua = (w == 0) ? 1 : 0;
  const int va = (w == 2) ? 1 : 2;
  uint64_t mb = spread3_s((uint32_t)K.i[I_G0V]) << va;
  const uint64_t ma0 = spread3_s((uint32_t)K.i[I_G0U]) << ua;
  for (int b = K.i[I_G0V]; b <= K.i[I_G1V]; b++, mb = morton_step(mb, va)) {
    const ColRow R = col_row(K, P, b);
    if (narrow) {
      int i0, i1;
      row_interval(K, R, P.unit, i0, i1);
      if (i0 > i1) continue;
      const uint64_t ma = (i0 == K.i[I_G0U]) ? ma0 : (spread3_s((uint32_t)i0) << ua);
      walk_row_cols(K, P, b, R, i0, i1, ma, mb, ua, act);
    } else {
      walk_row_cols(K, P, b, R, K.i[I_G0U], K.i[I_G1U], ma0, mb, ua, act);
    }
  }
}

// Global buffers the actions work on; each action binds its unit's segment.
struct ActArgs {
  uint32_t* table;    // all units' brick tables, concatenated
  uint8_t* pool8;     // touch path: 4096 bytes per slot
  ull* pool;          // 64 bit words per slot (written by the compaction)
  uint32_t* wprefix;  // per word: batch rank of its first voxel
  uint32_t* winner;   // per voxel of the batch: lowest covering triangle
  ull* ctr;
  uint64_t pool_cap;  // touch path slots
};

// Brick allocation without atomics: the touch pass stores 1 into the table entry of
// every brick a column's candidate voxels fall in; a scan over the table then numbers the
// flagged bricks in Morton order (table entry = slot + 1) and lists them.
struct TouchAction {
  static constexpr bool kTouch = true;
  // columns of a row per item of the small tier, per unit: with more than one, an item
  // walks its segment narrowed to the row's interval; atomic updates take one (the
  // lanes of a wave then update neighbouring cells, whose atomics share cache lines)
  __device__ static int seg_cols(const VoxParams&) { return 8; }
  // one-column items of a unit laid out over its rows' intervals only (see k_vox_tile)
  static constexpr bool kCompact = false;
  uint32_t* table;
  uint64_t brick0;  // first brick of the unit's table segment
  __device__ static TouchAction make(const ActArgs& A, const VoxParams& P) {
    return TouchAction{A.table + P.table_base, P.bstart >> kBrickBits};
  }
  __device__ __forceinline__ void touch(uint64_t brick_code) { table[brick_code - brick0] = 1u; }
  __device__ __forceinline__ void add(uint64_t, uint32_t) {}
  __device__ __forceinline__ void flush() {}
};

// Mark voxels with plain byte stores of 1: concurrent marks of one voxel all store the
// same value, so no atomic read-modify-write is needed however many triangles cover it
// (about three per voxel on a terrain).
//  * dense units (MarkDenseT): a byte per voxel of the unit's Morton range (the
//    reference's own layout, voxelizer.cpp:144, kept in HBM); the brick's table entry is
//    flagged with a plain store when the thread enters a brick. Only flagged bricks are
//    compacted and cleared afterwards, so the grid is never swept. Dense payload units
//    keep 4 bytes per voxel and resolve the winning triangle in the same walk (atomicMax
//    of ~triangle: voxelizer.cpp:263, the first triangle in partition order wins).
//  * touch units (MarkPoolAction): 4096 bytes per pool slot; the slot comes from the table
//    (touch pass + table scan), cached per thread.
// Both store through global (not generic) pointers.
typedef __attribute__((address_space(1))) uint8_t gbyte_t;
typedef __attribute__((address_space(1))) uint32_t gword_t;
// kWide: some units of the pass carry payload (4-byte cells; the unit's wide flag decides
// per voxel); otherwise every cell is a byte. kLaneOrder: the lanes of a wave hold
// non-decreasing triangle indices (the small tier: tile order, item order), and a lane
// checks its left neighbour before the atomic; not so in the large tier, whose list is in
// collection order.
template <bool kWide, bool kLaneOrder = true>
struct MarkDenseT {
  static constexpr bool kTouch = false;
  __device__ static int seg_cols(const VoxParams& P) {
    if (!kWide) return 16;
    if (kLaneOrder) return 1;  // the single-unit payload pass: its unit is wide
    return P.wide ? 1 : 16;
  }
  static constexpr bool kCompact = kWide;
  gword_t* table;
  gbyte_t* bytes;
  uint64_t base;  // code of the first voxel of the unit's first brick
  int wide;
  uint32_t cbrick = 0xFFFFFFFFu;
  __device__ static MarkDenseT make(const ActArgs& A, const VoxParams& P) {
    MarkDenseT m;
    m.table = (gword_t*)(A.table + P.table_base);
    m.bytes = (gbyte_t*)P.grid8;
    m.base = P.bstart;
    m.wide = kWide ? P.wide : 0;
    return m;
  }
  __device__ __forceinline__ void touch(uint64_t) {}
  __device__ __forceinline__ void add(uint64_t code, uint32_t tri) {
    const uint64_t rel = code - base;
    const uint32_t b = (uint32_t)(rel >> kBrickBits);
    if (kWide && wide) {
      // the lowest covering triangle as the largest ~tri (0 = unmarked). With kLaneOrder
      // a lane whose left neighbour marks the same cell in the same step holds a triangle
      // no lower than the neighbour's and leaves the atomic to it: on fine meshes
      // neighbouring triangles cover the same voxels
      uint32_t* cell = (uint32_t*)bytes + rel;
      bool dup = false;
      if (kLaneOrder) {
        const uint64_t addr = (uint64_t)cell;
        const uint64_t act = __ballot(1);
        const int ln = lane_id();
        const uint32_t plo = (uint32_t)__shfl_up((int)(uint32_t)addr, 1, kWave);
        const uint32_t phi = (uint32_t)__shfl_up((int)(uint32_t)(addr >> 32), 1, kWave);
        dup = ln > 0 && ((act >> (ln - 1)) & 1ull) && plo == (uint32_t)addr &&
              phi == (uint32_t)(addr >> 32);
      }
      if (!dup) atomicMax(cell, ~tri);
    } else
      bytes[rel] = 1;
    if (b != cbrick) {
      cbrick = b;
      table[b] = 1u;
    }
  }
  __device__ __forceinline__ void flush() {}
};
using MarkDenseAction = MarkDenseT<false>;
using MarkDenseWideAction = MarkDenseT<true>;
using MarkDenseWideBatchAction = MarkDenseT<true, false>;

struct MarkPoolAction {
  static constexpr bool kTouch = false;
  __device__ static int seg_cols(const VoxParams&) { return 8; }
  static constexpr bool kCompact = false;
  const gword_t* table;
  gbyte_t* bytes;  // the slot pool
  ull* ctr;
  uint64_t pool_cap;
  uint64_t base;  // code of the first voxel of the unit's first brick
  uint32_t cbrick = 0xFFFFFFFFu, cslot = 0;
  __device__ static MarkPoolAction make(const ActArgs& A, const VoxParams& P) {
    MarkPoolAction m;
    m.table = (const gword_t*)(A.table + P.table_base);
    m.bytes = (gbyte_t*)A.pool8;
    m.ctr = A.ctr;
    m.pool_cap = A.pool_cap;
    m.base = P.bstart;
    return m;
  }
  __device__ __forceinline__ void touch(uint64_t) {}
  __device__ __forceinline__ void add(uint64_t code, uint32_t) {
    const uint64_t rel = code - base;
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
    bytes[(uint64_t)(s - 1) * kBrickVox + (rel & (kBrickVox - 1))] = 1;
  }
  __device__ __forceinline__ void flush() {}
};

// Second payload pass: lowest triangle index per voxel rank.
struct WinnerAction {
  static constexpr bool kTouch = false;
  __device__ static int seg_cols(const VoxParams&) { return 1; }
  static constexpr bool kCompact = false;
  const uint32_t* table;
  const ull* pool;  // bit form of the bricks, written by the compaction
  const uint32_t* wprefix;
  uint32_t* winner;
  ull* ctr;
  uint64_t base;  // code of the first voxel of the unit's first brick
  __device__ static WinnerAction make(const ActArgs& A, const VoxParams& P) {
    return WinnerAction{A.table + P.table_base, A.pool, A.wprefix, A.winner, A.ctr, P.bstart};
  }
  __device__ __forceinline__ void touch(uint64_t) {}
  __device__ __forceinline__ void add(uint64_t code, uint32_t tri) {
    const uint64_t rel = code - base;
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

// last k with a[k] <= v, for a non-decreasing a[0..n) with a[0] <= v
template <typename T>
__device__ __forceinline__ uint32_t upper_index(const T* __restrict__ a, uint32_t n, T v) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (a[mid] <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Small tier, persistent: blocks take tiles of tt <= kTileTris consecutive triangles of one
// unit from a device work queue (tile_off[u] = first tile of unit u); tt is smaller for
// passes with few triangles, so that their tiles still spread over the CUs. For a tile, each thread builds
// its triangle's walk constants into LDS; a block scan lays out the (triangle, row
// segment) items, a segment being Act::seg_cols neighbouring columns of one row; then the
// threads share the items, so the work per lane is a few voxel tests regardless of
// triangle size. Triangles with more than small_cols columns
// (or non-slab triangles with more than small_cols * 8 tests) go to the column-parallel
// tier. Triangles with at most direct_cols columns skip the item layout: their own
// thread walks their columns from registers (the per-item lookup and constant loads
// would cost more than the walk). Units the pass does not apply to (skip mask)
// contribute no items.
constexpr int kTile = 256;      // threads per block
constexpr int kTileTris = 192;  // triangles per tile: 35 KB of walk constants in LDS, so
                                // that four blocks fit a CU
// Actions with kCompact (atomic payload marks, one column per item) lay out the items of a
// triangle of at most kCompactRows rows and 255 columns per row over its rows' intervals
// only (row_interval: the columns outside fail the projection test), 255 items at most;
// the rows' first columns and item offsets take 32 bytes of LDS per triangle, so their
// tiles hold kTileTrisCompact triangles (four blocks still fit a CU).
constexpr int kCompactRows = 16;
constexpr int kTileTrisCompact = 160;
constexpr uint32_t kQueues = 8;        // tile queues per persistent launch
constexpr uint32_t kQueueStride = 16;  // words between their counters (own cache lines)
constexpr uint32_t kQueueWords = kQueues * kQueueStride;
constexpr uint64_t kLargeWalkMax = 1u << 18;  // large triangles with stored walk constants
constexpr int kSkipDense = 1;   // units voxelized in a dense grid (touch pass)
constexpr int kSkipBinary = 2;  // geometry-only units (payload winner pass)

__device__ __forceinline__ bool skipped(const VoxParams& P, int skip) {
  return ((skip & kSkipDense) && P.dense) || ((skip & kSkipBinary) && P.stride == 9);
}

template <typename Act>
__global__ __launch_bounds__(kTile, 4) void k_vox_tile(const VoxParams* __restrict__ units,
                                                    uint32_t n_units,
                                                    const uint32_t* __restrict__ tile_off,
                                                    ActArgs A, ull* queue,
                                                    uint32_t* large_unit, uint32_t* large_tri,
                                                    uint32_t* large_cols, uint32_t* large_walk,
                                                    uint32_t walk_cap, uint32_t tt, int collect,
                                                    int skip) {
  constexpr bool kCompact = Act::kCompact;
  constexpr int kTT = kCompact ? kTileTrisCompact : kTileTris;
  __shared__ uint32_t sk[kWalkWords][kTT];
  __shared__ uint32_t soff[kTT + 1];
  __shared__ uint32_t wsum[kTile / kWave];
  __shared__ uint32_t s_tile;
  // compact layout, per triangle: items before each row, each row's first column - g0u,
  // and whether the triangle uses the layout
  __shared__ uint8_t s_rpre[kCompact ? kTT : 1][kCompactRows];
  __shared__ uint8_t s_ri0[kCompact ? kTT : 1][kCompactRows];
  __shared__ uint8_t s_cmp[kCompact ? kTT : 1];
  fill_spread_table();  // read after the first barrier of the tile loop
  ull* ctr = A.ctr;
  const uint32_t total_tiles = tile_off[n_units];
  const uint32_t tid = threadIdx.x;
  // kQueues interleaved tile classes (tile % kQueues), each with its own counter: one
  // counter's atomics serialize (~14 ns each), which bounds a single queue at ~70 M
  // tiles/s. A block starts on class blockIdx % kQueues and moves on when it is empty.
  uint32_t cls = blockIdx.x % kQueues, tried = 0;
  for (;;) {
    if (tid == 0) {
      uint32_t tl = total_tiles;
      while (tried < kQueues) {
        const uint32_t k = (uint32_t)atomicAdd(queue + kQueueStride * cls, 1ull);
        const uint32_t cand = cls + kQueues * k;
        if (cand < total_tiles) {
          tl = cand;
          break;
        }
        cls = (cls + 1) % kQueues;
        tried++;
      }
      s_tile = tl;
    }
    __syncthreads();
    // the block's tile in a scalar register: the unit search and the unit's parameters
    // then run on scalar loads
    const uint32_t tile = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_tile);
    if (tile >= total_tiles) break;
    const uint32_t u = upper_index<uint32_t>(tile_off, n_units + 1, tile);
    const VoxParams P = units[u];
    const bool active = !skipped(P, skip);
    const uint32_t i = (tile - tile_off[u]) * tt + tid;
    uint32_t items = 0;
    if (active && tid < tt && i < P.n) {
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
          large_unit[idx] = u;
          large_tri[idx] = t;
          large_cols[idx] = (uint32_t)ncols;
          if (idx < walk_cap) {  // its walk constants, for the column threads
            uint32_t* dst = large_walk + idx * kWalkWords;
#pragma unroll
            for (int f = 0; f < kWalkF; f++) dst[f] = __float_as_uint(K.f[f]);
#pragma unroll
            for (int f = 0; f < kWalkI; f++) dst[kWalkF + f] = (uint32_t)K.i[f];
          }
        }
      } else if (ncols <= (uint64_t)P.direct_cols) {
        // few columns: this thread walks them now, from registers
        Act act = Act::make(A, P);
        walk_all_columns(K, P, act, Act::seg_cols(P) > 1);
        act.flush();
      } else {
        // items: the rows' segments of seg columns
        const uint32_t seg = (uint32_t)Act::seg_cols(P);
        const uint32_t eu = (uint32_t)(K.i[I_G1U] - K.i[I_G0U] + 1);
        const uint32_t nrows = (uint32_t)(K.i[I_G1V] - K.i[I_G0V] + 1);
        items = nrows * ((eu + seg - 1) / seg);
        if constexpr (kCompact) {
          // one-column items over the rows' intervals only
          bool cmp = false;
          if (seg == 1 && nrows <= (uint32_t)kCompactRows && eu <= 255u) {
            uint32_t cum = 0;
            for (uint32_t r = 0; r < nrows; r++) {
              const ColRow R = col_row(K, P, K.i[I_G0V] + (int)r);
              int i0, i1;
              row_interval(K, R, P.unit, i0, i1);
              s_rpre[tid][r] = (uint8_t)min(cum, 255u);
              s_ri0[tid][r] = (uint8_t)min(i0 - K.i[I_G0U], 255);
              if (i1 >= i0) cum += (uint32_t)(i1 - i0 + 1);
            }
            if (cum <= 255u) {
              cmp = true;
              items = cum;
            }
          }
          s_cmp[tid] = cmp ? 1 : 0;
        }
#pragma unroll
        for (int f = 0; f < kWalkF; f++) sk[f][tid] = __float_as_uint(K.f[f]);
#pragma unroll
        for (int f = 0; f < kWalkI; f++) sk[kWalkF + f][tid] = (uint32_t)K.i[f];
      }
    }
    uint32_t total;
    const uint32_t inc = block_incl_scan<uint32_t, kTile>(items, wsum, total);
    if (tid < tt) soff[tid] = inc - items;
    if (tid == tt - 1) soff[tt] = total;
    __syncthreads();
    Act act = Act::make(A, P);
    for (uint32_t it = tid; it < total; it += kTile) {
      // owner: last j with soff[j] <= it (triangles without items have empty ranges)
      const uint32_t lo = upper_index<uint32_t>(soff, tt, it);
      Walk K;
#pragma unroll
      for (int f = 0; f < kWalkF; f++) K.f[f] = __uint_as_float(sk[f][lo]);
#pragma unroll
      for (int f = 0; f < kWalkI; f++) K.i[f] = (int)sk[kWalkF + f][lo];
      const uint32_t local = it - soff[lo];
      const uint32_t eu = (uint32_t)(K.i[I_G1U] - K.i[I_G0U] + 1);
      const int seg = Act::seg_cols(P);  // the tile's unit: uniform
      if (kCompact && s_cmp[kCompact ? lo : 0]) {
        // row: the last r with s_rpre[r] <= local among the triangle's rows (an empty
        // row shares its offset with the next one, which is the one taken)
        const uint32_t nrows = (uint32_t)(K.i[I_G1V] - K.i[I_G0V] + 1);
        uint32_t r0 = 0, r1 = nrows - 1;
        while (r0 < r1) {
          const uint32_t mid = (r0 + r1 + 1) >> 1;
          if (s_rpre[lo][mid] <= local)
            r0 = mid;
          else
            r1 = mid - 1;
        }
        const int a = K.i[I_G0U] + (int)s_ri0[lo][r0] + (int)(local - s_rpre[lo][r0]);
        walk_row_span<false>(K, P, K.i[I_G0V] + (int)r0, a, a, act);
      } else if (seg > 1) {
        const uint32_t nseg = (eu + seg - 1) / seg;
        const int a = K.i[I_G0U] + (int)(local % nseg) * seg;
        const int b = K.i[I_G0V] + (int)(local / nseg);
        walk_row_span<true>(K, P, b, a, min(a + seg - 1, K.i[I_G1U]), act);
      } else {
        const int a = K.i[I_G0U] + (int)(local % eu);
        const int b = K.i[I_G0V] + (int)(local / eu);
        walk_row_span<false>(K, P, b, a, a, act);
      }
    }
    act.flush();
    __syncthreads();  // LDS (setups, offsets, s_tile) is reused by the next tile
  }
}

// Column-parallel tier: one thread per column of the deferred triangles; the counts live
// on the device (ctr[kCtrLargeCount] triangles, ctr[kCtrCols] columns), the grid strides.
// walk constants stored as 44 words (kWalkF floats, then kWalkI ints)
__device__ __forceinline__ void load_walk(const uint32_t* __restrict__ src, Walk& K) {
#pragma unroll
  for (int f = 0; f < kWalkF; f++) K.f[f] = __uint_as_float(src[f]);
#pragma unroll
  for (int f = 0; f < kWalkI; f++) K.i[f] = (int)src[kWalkF + f];
}

template <typename Act>
__global__ __launch_bounds__(256) void k_vox_large(const VoxParams* __restrict__ units, ActArgs A,
                                                   const uint32_t* __restrict__ large_unit,
                                                   const uint32_t* __restrict__ large_tri,
                                                   const uint64_t* __restrict__ large_off,
                                                   const uint32_t* __restrict__ large_walk,
                                                   uint32_t walk_cap, int skip) {
  fill_spread_table();
  __syncthreads();
  __shared__ uint32_t s_j0;
  const ull* ctr = A.ctr;
  const uint32_t n_large = (uint32_t)ctr[kCtrLargeCount];
  const uint64_t total = ctr[kCtrCols];
  // each block walks one contiguous range of the columns, 256 at a time: the owner of its
  // first column comes from one binary search, and a thread's owner only moves forward
  // (a triangle here has more than small_cols columns), one offset compare per step
  const uint64_t per = (total + gridDim.x - 1) / gridDim.x;
  const uint64_t beg = (uint64_t)blockIdx.x * per;
  const uint64_t end = min(total, beg + per);
  if (beg >= end) return;  // uniform per block
  if (threadIdx.x == 0) s_j0 = upper_index<uint64_t>(large_off, n_large, beg);
  __syncthreads();
  uint32_t j = s_j0;
  for (uint64_t g = beg + threadIdx.x; g < end; g += blockDim.x) {
    while (j + 1 < n_large && large_off[j + 1] <= g) j++;
    // the lanes of a wave nearly always share their triangle: then its unit parameters
    // and walk constants come in with scalar loads (one per 16 words) instead of some 60
    // vector loads per column, which bound this kernel
    // (a wave that straddles two triangles sets its lanes' walks up again)
    const uint32_t jf = (uint32_t)__builtin_amdgcn_readfirstlane((int)j);
    VoxParams P;
    Walk K;
    uint32_t t;
    uint64_t local;
    bool stored = false;
    if (__all(j == jf)) {
      P = units[large_unit[jf]];
      t = large_tri[jf];
      local = g - large_off[jf];
      if (jf < walk_cap) {
        load_walk(large_walk + (uint64_t)jf * kWalkWords, K);
        stored = true;
      }
    } else {
      P = units[large_unit[j]];
      t = large_tri[j];
      local = g - large_off[j];
    }
    if (skipped(P, skip)) continue;
    if (!stored) {
      Setup S;
      setup_tri(P.tris + (uint64_t)t * P.stride, P, S);
      make_walk(S, P, t, K);
    }
    const uint32_t eu = (uint32_t)(K.i[I_G1U] - K.i[I_G0U] + 1);
    int a, b;
    if (local <= 0xFFFFFFFFull) {  // 32-bit division (a 64-bit one is a long routine)
      a = K.i[I_G0U] + (int)((uint32_t)local % eu);
      b = K.i[I_G0V] + (int)((uint32_t)local / eu);
    } else {
      a = K.i[I_G0U] + (int)(local % eu);
      b = K.i[I_G0V] + (int)(local / eu);
    }
    Act act = Act::make(A, P);
    walk_row_span<false>(K, P, b, a, a, act);
    act.flush();
  }
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
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t v, bool last) const {
    if (v) {
      occ[ex] = (uint32_t)i;
      table[i] = ex + 1u;
    }
    if (last) ctr[kCtrNumOcc] = (ull)(ex + v);
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

// Pack each flagged brick's bytes into 64 bit words (lane l = voxels 64l..64l+63) at the
// brick's rank, count its voxels, and zero the bytes that were set. The bytes live in the
// unit's dense grid (at the brick's index) or in the slot pool (at its rank).
// ubase[u] = first table entry of unit u (n_units + 1 entries).
// Dense passes flag only bricks that hold voxels, so the rank is the brick's place in the
// brick stream: there 'pool' is the stream's word array, bcode (non-null) receives the
// tagged brick codes and the table entries are cleared here.
// ONE: a single unit (its parameters are uniform, read with scalar loads).

// Brick wv (table entry ti, unit u): its tagged code, and its table entry cleared.
__device__ __forceinline__ void brick_head(uint64_t wv, uint64_t ti, uint32_t u,
                                           const VoxParams* __restrict__ units,
                                           const uint64_t* __restrict__ ubase,
                                           uint64_t* __restrict__ bcode,
                                           uint32_t* __restrict__ table) {
  if (bcode && lane_id() == 0) {
    bcode[wv] = (units[u].bstart + ((ti - ubase[u]) << kBrickBits)) |
                ((uint64_t)units[u].job << kJobShift);
    table[ti] = 0u;
  }
}
// The lane's 16-byte pieces of a brick's bytes: load q of lane l is the 16 bytes at
// q * 1024 + 16 l (each load of the wave one contiguous KiB).
__device__ __forceinline__ uint4* brick_bytes(uint64_t wv, uint64_t ti, uint32_t u,
                                              const VoxParams* __restrict__ units,
                                              const uint64_t* __restrict__ ubase,
                                              uint8_t* __restrict__ pool8) {
  uint8_t* base = units[u].dense ? units[u].grid8 + (ti - ubase[u]) * kBrickVox
                                 : pool8 + wv * kBrickVox;
  return (uint4*)base + lane_id();
}
// The lane's occupancy word from the loaded pieces (piece q of lane l holds bits
// 16 (l % 4) .. + 15 of word 16 q + l / 4); pieces that held marks are zeroed.
__device__ __forceinline__ ull pack_brick_bytes(uint4* src, const uint4 (&v)[4]) {
  ull parts = 0;  // 16 bits per load
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const uint32_t w4[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
    for (int d = 0; d < 4; d++) {
      const uint32_t x = w4[d];  // four bytes, each 0 or 1
      const uint32_t nib = (x & 1u) | ((x >> 7) & 2u) | ((x >> 14) & 4u) | ((x >> 21) & 8u);
      parts |= (ull)nib << (16 * q + 4 * d);
    }
    if (v[q].x | v[q].y | v[q].z | v[q].w) src[q * kWave] = make_uint4(0u, 0u, 0u, 0u);
  }
  // word L = 16 q + r gathers the parts of load q from lanes 4 r .. 4 r + 3
  const int L = lane_id(), q = L >> 4, r = L & 15;
  ull m = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)parts, 4 * r + k, kWave);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(parts >> 32), 4 * r + k, kWave);
    const ull p = ((ull)hi << 32) | lo;
    m |= ((p >> (16 * q)) & 0xFFFFull) << (16 * k);
  }
  return m;
}
// Brick wv's word array, its node shape (dense passes) and its voxel count.
__device__ __forceinline__ void brick_tail(uint64_t wv, ull m, ull* __restrict__ pool,
                                           uint64_t* __restrict__ bcount,
                                           const uint64_t* __restrict__ bcode,
                                           uint32_t* __restrict__ bshape) {
  pool[wv * kBrickWords + lane_id()] = m;
  if (bcode) {  // the brick node's subtree size and child mask, for the build
    const BrickShape B = brick_shape_of(m);
    if (lane_id() == 0) bshape[wv] = brick_shape_word(B);
  }
  const uint32_t c = (uint32_t)__popcll(m);
  const uint32_t tot = wave_incl_scan<uint32_t>(c);
  // voxel count, plus a flag in bit kNonEmptyShift: one scan gives every brick its
  // voxel offset and its rank among the non-empty bricks (the touch pass flags a
  // superset of the bricks that end up holding voxels)
  if (lane_id() == kWave - 1) bcount[wv] = tot + (tot ? (1ull << kNonEmptyShift) : 0ull);
}

template <bool ONE>
__device__ __forceinline__ void brick_count_body(const uint32_t* __restrict__ occ,
                                                 uint64_t n_occ,
                                                 const VoxParams* __restrict__ units,
                                                 const uint64_t* __restrict__ ubase,
                                                 uint32_t n_units, uint8_t* __restrict__ pool8,
                                                 ull* __restrict__ pool,
                                                 uint64_t* __restrict__ bcount,
                                                 uint64_t* __restrict__ bcode,
                                                 uint32_t* __restrict__ bshape,
                                                 uint32_t* __restrict__ table) {
  for (uint64_t wv = wave_global(); wv < n_occ; wv += waves_total()) {
    // the brick is the wave's: its table index in a scalar register, so that the unit
    // search below runs on scalar loads
    const uint64_t ti = (uint32_t)__builtin_amdgcn_readfirstlane((int)occ[wv]);
    const uint32_t u = ONE ? 0u : upper_index<uint64_t>(ubase, n_units + 1, ti);
    brick_head(wv, ti, u, units, ubase, bcode, table);
    ull m = 0;
    if (units[u].wide) {
      // 4-byte cells (~winner, 0 = unmarked): kept for k_brick_emit, which clears them
      const uint4* src = (const uint4*)((const uint32_t*)units[u].grid8 +
                                        (ti - ubase[u]) * kBrickVox + lane_id() * 64);
      // four loads at a time: few registers, so that many waves stay resident
#pragma unroll 1
      for (int g = 0; g < 16; g += 4) {
        uint4 v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = src[g + q];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint32_t nib = (v[q].x != 0u ? 1u : 0u) | (v[q].y != 0u ? 2u : 0u) |
                               (v[q].z != 0u ? 4u : 0u) | (v[q].w != 0u ? 8u : 0u);
          m |= (ull)nib << (4 * (g + q));
        }
      }
    } else {
      uint4* src = brick_bytes(wv, ti, u, units, ubase, pool8);
      uint4 v[4];
#pragma unroll
      for (int q = 0; q < 4; q++) v[q] = src[q * kWave];
      m = pack_brick_bytes(src, v);
    }
    brick_tail(wv, m, pool, bcount, bcode, bshape);
  }
}

__global__ __launch_bounds__(256) void k_brick_count(const uint32_t* __restrict__ occ,
                                                     const ull* __restrict__ ctr, uint64_t cap,
                                                     const VoxParams* __restrict__ units,
                                                     const uint64_t* __restrict__ ubase,
                                                     uint32_t n_units, uint8_t* __restrict__ pool8,
                                                     ull* __restrict__ pool,
                                                     uint64_t* __restrict__ bcount,
                                                     uint64_t* __restrict__ bcode,
                                                     uint32_t* __restrict__ bshape,
                                                     uint32_t* __restrict__ table) {
  // more flagged bricks than slots (touch path): the pass is redone with a larger pool
  const uint64_t n_occ = min((uint64_t)ctr[kCtrNumOcc], cap);
  if (n_units == 1)
    brick_count_body<true>(occ, n_occ, units, ubase, n_units, pool8, pool, bcount, bcode,
                           bshape, table);
  else
    brick_count_body<false>(occ, n_occ, units, ubase, n_units, pool8, pool, bcount, bcode,
                            bshape, table);
}

struct CountStore {
  uint64_t* off;
  ull* total;
  __device__ void operator()(uint64_t i, uint64_t ex, uint64_t v, bool last) const {
    off[i] = ex;
    if (last) *total = (ull)(ex + v);
  }
};
// Brick counts of a dense pass: also each brick's first voxel in the stream (s0 + offset).
struct BrickOffStore {
  uint64_t* off;
  ull* total;
  uint64_t* bvoff;
  uint64_t s0;
  __device__ void operator()(uint64_t i, uint64_t ex, uint64_t v, bool last) const {
    off[i] = ex;
    bvoff[i] = s0 + (ex & kVoxMask);
    if (last) *total = (ull)(ex + v);
  }
};
struct ColLoad {
  const uint32_t* a;
  __device__ uint64_t operator()(uint64_t i) const { return a[i]; }
};

// The brick stream entry of every non-empty flagged brick and, with kCodes, for payload
// units the sorted codes of its voxels, tagged with the unit's job in bits kJobShift+
// (and the payload winners of dense payload units). stage_cells: the launch gives each
// wave kBrickVox words of dynamic LDS, into which a dense payload brick's 4-byte cells
// are copied with coalesced loads (and cleared in the grid) before its voxels read their
// winners from there, instead of one dependent global load per voxel.
template <bool kCodes>
__global__ __launch_bounds__(256) void k_brick_emit(const uint32_t* __restrict__ occ,
                                                    const ull* __restrict__ ctr,
                                                    const VoxParams* __restrict__ units,
                                                    const uint64_t* __restrict__ ubase,
                                                    uint32_t n_units,
                                                    const ull* __restrict__ pool,
                                                    const uint64_t* __restrict__ boff,
                                                    uint64_t* __restrict__ codes,
                                                    uint32_t* __restrict__ wprefix,
                                                    uint64_t s0, uint64_t* __restrict__ bwords,
                                                    uint64_t* __restrict__ bcode,
                                                    uint64_t* __restrict__ bvoff,
                                                    uint32_t* __restrict__ bshape,
                                                    uint32_t* __restrict__ winner,
                                                    int stage_cells, int write_stream) {
  extern __shared__ uint32_t s_cells[];  // stage_cells: kBrickVox per wave
  const uint64_t n_occ = ctr[kCtrNumOcc];
  for (uint64_t wv = wave_global(); wv < n_occ; wv += waves_total()) {
    const uint64_t ti = (uint32_t)__builtin_amdgcn_readfirstlane((int)occ[wv]);  // wave's brick
    const uint32_t u = upper_index<uint64_t>(ubase, n_units + 1, ti);
    // a dense pass has its brick stream already (k_brick_count): only the payload units'
    // bricks have work left
    if (!write_stream && units[u].stride == 9) continue;
    const uint64_t slot = wv * kBrickWords + lane_id();
    ull word = pool[slot];
    const uint64_t bo = boff[wv];
    const uint64_t tag = (uint64_t)units[u].job << kJobShift;
    const uint64_t b0 = units[u].bstart + ((ti - ubase[u]) << kBrickBits);
    // the brick stream (non-empty bricks, appended in stream order)
    if (write_stream && __ballot(word != 0ull)) {
      const uint64_t sb = bo >> kNonEmptyShift;
      bwords[sb * kBrickWords + lane_id()] = word;
      const BrickShape B = brick_shape_of(word);
      if (lane_id() == 0) {
        bcode[sb] = b0 | tag;
        bvoff[sb] = s0 + (bo & kVoxMask);
        bshape[sb] = brick_shape_word(B);
      }
    }
    if constexpr (!kCodes) continue;
    if (units[u].stride == 9) continue;  // geometry-only unit: codes written on demand
    const uint32_t c = (uint32_t)__popcll(word);
    const uint32_t inc = wave_incl_scan<uint32_t>(c);
    uint64_t pos = (bo & kVoxMask) + (inc - c);
    if (wprefix) wprefix[slot] = (uint32_t)pos;
    const uint64_t cb = (b0 + ((uint64_t)lane_id() << 6)) | tag;
    // dense payload unit: each voxel's winner from its cell, which is cleared
    if (units[u].wide && stage_cells) {
      uint32_t* sc = s_cells + (threadIdx.x / kWave) * kBrickVox;
      uint4* src = (uint4*)((uint32_t*)units[u].grid8 + (ti - ubase[u]) * kBrickVox);
      uint4 v[kBrickVox / 4 / kWave];
      // uint4 r * 64 + l holds cells of word 4 r + l / 16: the cells of an empty word are
      // all zero, so only those of non-empty words are read
      const uint64_t ne = __ballot(word != 0ull);
#pragma unroll
      for (int r = 0; r < kBrickVox / 4 / kWave; r++)
        v[r] = ((ne >> (4 * r + (lane_id() >> 4))) & 1ull) ? src[r * kWave + lane_id()]
                                                           : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
      for (int r = 0; r < kBrickVox / 4 / kWave; r++) {
        ((uint4*)sc)[r * kWave + lane_id()] = v[r];
        if (v[r].x | v[r].y | v[r].z | v[r].w) src[r * kWave + lane_id()] = make_uint4(0u, 0u, 0u, 0u);
      }
      wave_lds_sync();
      const uint32_t* mine = sc + lane_id() * 64;  // the cells of the lane's word
      while (word) {
        const int bit = __ffsll(word) - 1;
        winner[pos] = ~mine[bit];
        codes[pos++] = cb + (uint64_t)bit;
        word &= word - 1;
      }
      wave_lds_sync();  // the next brick's cells overwrite these
      continue;
    }
    uint32_t* cells = units[u].wide ? (uint32_t*)units[u].grid8 + (ti - ubase[u]) * kBrickVox +
                                          lane_id() * 64
                                    : nullptr;
    while (word) {
      const int bit = __ffsll(word) - 1;
      if (cells) {
        winner[pos] = ~cells[bit];
        cells[bit] = 0u;
      }
      codes[pos++] = cb + (uint64_t)bit;
      word &= word - 1;
    }
  }
}

// Codes of the voxels of bricks [0, nb) of the stream, from their occupancy words: one
// wave per brick, lane l = word l, written at the brick's stream offset.
__global__ __launch_bounds__(256) void k_bricks_codes(const uint64_t* __restrict__ bwords,
                                                      const uint64_t* __restrict__ bcode,
                                                      const uint64_t* __restrict__ bvoff,
                                                      uint64_t nb, uint64_t* __restrict__ codes) {
  for (uint64_t b = wave_global(); b < nb; b += waves_total()) {
    ull word = bwords[b * kBrickWords + lane_id()];
    const uint32_t c = (uint32_t)__popcll(word);
    const uint32_t inc = wave_incl_scan<uint32_t>(c);
    uint64_t pos = bvoff[b] + (inc - c);
    const uint64_t cb = bcode[b] + ((uint64_t)lane_id() << 6);
    while (word) {
      const int bit = __ffsll(word) - 1;
      codes[pos++] = cb + (uint64_t)bit;
      word &= word - 1;
    }
  }
}

// restore the all-zero state of the table for the flagged bricks (their bytes were
// cleared by k_brick_count)
__global__ __launch_bounds__(256) void k_brick_clear(const uint32_t* __restrict__ occ,
                                                     const ull* __restrict__ ctr,
                                                     uint32_t* __restrict__ table) {
  const uint64_t n_occ = ctr[kCtrNumOcc];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_occ; i += stride)
    table[occ[i]] = 0u;
}

// First voxel (batch rank) of each unit: the offset of its first flagged brick.
__global__ void k_unit_voxels(const uint32_t* __restrict__ occ, const ull* __restrict__ ctr,
                              const uint64_t* __restrict__ ubase, uint32_t n_units,
                              const uint64_t* __restrict__ boff, uint64_t* __restrict__ uvoff) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u > n_units) return;
  const uint64_t n_occ = ctr[kCtrNumOcc];
  const uint64_t V = ctr[kCtrNumVox] & kVoxMask;
  if (u == n_units) {
    uvoff[u] = V;
    return;
  }
  // first occupied brick with table index >= ubase[u]
  uint64_t lo = 0, hi = n_occ;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if ((uint64_t)occ[mid] < ubase[u])
      lo = mid + 1;
    else
      hi = mid;
  }
  uvoff[u] = (lo < n_occ) ? (boff[lo] & kVoxMask) : V;
}

// ---------------------------------------------------------------- payload records
// BarycentricCoords.h:4-32 for the winning triangle; colour modes of main.cpp:373-382.
// Voxel i of the pass (stream index s0 + i) belongs to the unit whose voxel range holds
// it, and to that unit's job; its record goes to dbase[job] + 1 + (s0 + i - voff[job]).
// Voxels of geometry-only jobs have no record here.
// Threads cover the payload units' voxels only: payload voxel k lies in the payload unit
// j with prange[j] <= k < prange[j + 1] (a block searches for its first voxel, a thread
// steps forward over the rare unit boundaries), unit punit[j], whose voxels start at pass
// voxel pstart[j].
__global__ __launch_bounds__(256) void k_payload_records(
    const JobDev* __restrict__ jobs, const VoxParams* __restrict__ units,
    const uint64_t* __restrict__ prange, const uint64_t* __restrict__ pstart,
    const uint32_t* __restrict__ punit, uint32_t n_pu, const uint64_t* __restrict__ codes,
    uint64_t s0, const uint32_t* __restrict__ winner, uint64_t n_pv, float* __restrict__ out,
    ull* ctr) {
  __shared__ uint32_t s_j0;
  const uint64_t b0 = (uint64_t)blockIdx.x * blockDim.x;
  const uint64_t k = b0 + threadIdx.x;
  if (threadIdx.x == 0) s_j0 = upper_index<uint64_t>(prange, n_pu + 1, b0);
  __syncthreads();
  if (k >= n_pv) return;
  uint32_t j = s_j0;
  while (j + 1 < n_pu && prange[j + 1] <= k) j++;
  const uint32_t u = punit[j];
  const uint64_t i = pstart[j] + (k - prange[j]);  // voxel i of the pass
  const uint64_t tagged = codes[s0 + i];
  const JobDev& J = jobs[units[u].job];
  if (!J.payload) return;
  const uint64_t code = tagged & kCodeMask;
  const uint32_t t = winner[i];
  float col0, col1, col2, nr0, nr1, nr2;
  if (t == 0xFFFFFFFFu) {  // cannot happen: both passes evaluate the same predicates
    atomicAdd(&ctr[kCtrError], 1ull);
    col0 = col1 = col2 = nr0 = nr1 = nr2 = 0.0f;
  } else {
    const float* p = J.tris + (uint64_t)t * 21;
    const float v0x = p[0], v0y = p[1], v0z = p[2];
    const float v1x = p[3], v1y = p[4], v1z = p[5];
    const float v2x = p[6], v2y = p[7], v2z = p[8];
    nr0 = p[9];
    nr1 = p[10];
    nr2 = p[11];
    uint32_t x, y, z;
    unmorton3(code, x, y, z);
    const int color_mode = J.color_mode;
    const float unit_div = J.unit_div;
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
      const float g = (float)J.gridsize;
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
  float* r = out + (J.dbase + 1 + (s0 + i - J.voff)) * 8;
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
  hitbox[t] = part_hitbox_of(tris + t * stride, cube_lo, cube_hi, K);
}

// out[k] = the record of triangle list[k] (list == nullptr: triangle k), stride floats each
__global__ __launch_bounds__(256) void k_gather_records(const float* __restrict__ tris, int stride,
                                                        const uint32_t* __restrict__ list,
                                                        uint64_t n, float* __restrict__ out) {
  const uint64_t total = n * (uint64_t)stride;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = e / (uint64_t)stride, f = e - k * (uint64_t)stride;
    const uint64_t t = list ? list[k] : k;
    out[e] = tris[t * stride + f];
  }
}

hipError_t gather_records(Context& c, const float* tris, int stride, const uint32_t* list,
                          uint64_t n, float* out) {
  if (!n) return hipSuccess;
  const uint64_t total = n * (uint64_t)stride;
  SVO_LAUNCH(k_gather_records, dim3((unsigned)std::min<uint64_t>((total + 255) / 256, 16384)),
             dim3(256), 0, c.stream, tris, stride, list, n, out);
  return hipGetLastError();
}

struct PartStore {
  uint32_t* list;
  ull* total;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t v, bool last) const {
    if (v) list[ex] = (uint32_t)i;
    if (last) *total = (ull)(ex + v);
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
// two small ranges zeroed by one launch
__global__ void k_zero2_u64(uint64_t* a, uint64_t na, uint64_t* b, uint64_t nb) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na + nb; i += stride) {
    if (i < na)
      a[i] = 0;
    else
      b[i - na] = 0;
  }
}
static unsigned fill_grid(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 8192));
}
hipError_t fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  SVO_LAUNCH(k_fill_u32, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}
hipError_t fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  SVO_LAUNCH(k_fill_u64, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------- host drivers
#define SVO_TRY(x)                        \
  do {                                    \
    hipError_t _e = (x);                  \
    if (_e != hipSuccess) return _e;      \
  } while (0)

hipError_t upload_async(Context& c, void* dst, const void* src, size_t bytes) {
  if (!bytes) return hipSuccess;
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (c.stage_used + need > c.stage_cap) {
    // every copy from the staging buffer has completed after this
    SVO_TRY(stream_sync(c.stream));
    c.stage_used = 0;
    if (need > c.stage_cap) {
      if (c.stage) (void)hipHostFree(c.stage);
      c.stage = nullptr;
      c.stage_cap = 0;
      const size_t cap = std::max<size_t>(need, 4u << 20);
      SVO_TRY(hipHostMalloc(&c.stage, cap, hipHostMallocDefault));
      c.stage_cap = cap;
    }
  }
  void* h = (char*)c.stage + c.stage_used;
  memcpy(h, src, bytes);
  c.stage_used += need;
  return hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, c.stream);
}

// Small read-backs: one kernel copies every request into the pinned host buffer (mapped
// into the device's address space), instead of one DMA copy per request.
constexpr int kGatherMax = 4;
struct GatherArgs {
  const uint64_t* src[kGatherMax];
  uint64_t* dst[kGatherMax];
  uint32_t words[kGatherMax];
  int n;
};
// Then it raises *flag = seq (after a system-scope fence), which the host polls: a
// shorter wait than a stream synchronization.
__global__ __launch_bounds__(256) void k_gather_host(GatherArgs g, volatile uint32_t* flag,
                                                     uint32_t seq) {
  for (int k = 0; k < g.n; k++)
    for (uint32_t i = threadIdx.x; i < g.words[k]; i += blockDim.x) g.dst[k][i] = g.src[k][i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) *flag = seq;
}

hipError_t read_back(Context& c, const ReadReq* reqs, int n) {
  t_syncs++;  // one host wait, whichever way it is done
  size_t total = 0;
  bool small = n <= kGatherMax;
  for (int k = 0; k < n; k++) {
    total += (reqs[k].bytes + 255) & ~(size_t)255;
    small = small && reqs[k].bytes % 8 == 0 && (uintptr_t)reqs[k].src % 8 == 0;
  }
  small = small && total <= (64u << 10);
  total += 256;  // the completion flag
  if (total > c.hread_cap) {
    SVO_TRY(hipStreamSynchronize(c.stream));
    if (c.hread) (void)hipHostFree(c.hread);
    c.hread = nullptr;
    c.hread_cap = 0;
    const size_t cap = std::max<size_t>(total, 1u << 20);
    SVO_TRY(hipHostMalloc(&c.hread, cap, hipHostMallocDefault));
    c.hread_cap = cap;
  }
  const size_t base = 256;
  size_t off = base;
  if (small) {
    GatherArgs g{};
    g.n = n;
    for (int k = 0; k < n; k++) {
      g.src[k] = (const uint64_t*)reqs[k].src;
      g.dst[k] = (uint64_t*)((char*)c.hread + off);
      g.words[k] = (uint32_t)(reqs[k].bytes / 8);
      off += (reqs[k].bytes + 255) & ~(size_t)255;
    }
    volatile uint32_t* flag = (volatile uint32_t*)c.hread;
    const uint32_t seq = ++c.hread_seq;
    SVO_LAUNCH(k_gather_host, dim3(1), dim3(256), 0, c.stream, g, flag, seq);
    SVO_TRY(hipGetLastError());
    // poll the flag for a short while (the usual case: the stream is almost done), then
    // fall back to a blocking synchronization (long kernels ahead, or a failed launch)
    bool seen = false;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t k = 0;; k++) {
      if (*flag == seq) {
        seen = true;
        break;
      }
      if ((k & 255) == 255 &&
          std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(2000))
        break;
    }
    if (seen)
      std::atomic_thread_fence(std::memory_order_acquire);
    else
      SVO_TRY(hipStreamSynchronize(c.stream));
  } else {
    for (int k = 0; k < n; k++) {
      if (reqs[k].bytes)
        SVO_TRY(hipMemcpyAsync((char*)c.hread + off, reqs[k].src, reqs[k].bytes,
                               hipMemcpyDeviceToHost, c.stream));
      off += (reqs[k].bytes + 255) & ~(size_t)255;
    }
    SVO_TRY(hipStreamSynchronize(c.stream));
  }
  off = base;
  for (int k = 0; k < n; k++) {
    memcpy(reqs[k].dst, (char*)c.hread + off, reqs[k].bytes);
    off += (reqs[k].bytes + 255) & ~(size_t)255;
  }
  return hipSuccess;
}

static hipError_t ensure_scan_scratch(Context& c, uint64_t n, size_t elem) {
  return ensure(c.scan_scratch, scan_scratch_elems(n) * elem + 256);
}

hipError_t partition_job(Context& c, uint32_t jid, float unitlength, uint64_t gridsize,
                         uint64_t n_parts) {
  Job& J = c.jobs[jid];
  J.n_parts = n_parts;
  J.part_off.assign(n_parts, 0);
  J.part_cnt.assign(n_parts, 0);
  if (n_parts <= 1) {
    J.part_cnt[0] = J.n_tris;
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
  std::vector<float> lohi(2 * 256, 0.0f);
  for (int i = 0; i < K; i++) {
    lohi[i] = (float)(uint32_t)(i * side) * unitlength;
    lohi[256 + i] = (float)(uint32_t)((i + 1) * side) * unitlength;
  }
  const uint64_t n = J.n_tris;
  const int stride = J.payload ? 21 : 9;
  SVO_TRY(ensure(c.hitbox, n * 8 + 2 * 256 * sizeof(float) + 256));
  float* d_lo = (float*)((char*)c.hitbox.p + ((n * 8 + 255) & ~255ull));
  float* d_hi = d_lo + 256;
  SVO_TRY(upload_async(c, d_lo, lohi.data(), 2 * 256 * sizeof(float)));
  if (n) {
    ktimer_begin(c, "part_hitbox");
    SVO_LAUNCH(k_part_hitbox, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream,
                       J.tris, stride, n, d_lo, d_hi, K, c.hitbox.as<uint64_t>());
    ktimer_end(c);
    SVO_TRY(hipGetLastError());
  }
  // one order-preserving flag scan per partition; lists concatenated in partition order
  SVO_TRY(ensure_scan_scratch(c, n, 4));
  ull* d_tot = c.counters.as<ull>() + kCtrPartTotal;
  for (uint64_t j = 0; j < n_parts; j++) {
    // list capacity: grow if the previous partitions filled it
    if ((c.part_used + n) * 4 > c.part_lists.cap)
      SVO_TRY(ensure_keep(c.part_lists, (c.part_used + n) * 4 * 2 + 256, c.part_used * 4,
                          c.stream));
    SVO_TRY(fill_u64((uint64_t*)d_tot, 0, 1, c.stream));
    uint32_t cx, cy, cz;
    unmorton3(j, cx, cy, cz);
    if (n) {
      PartLoad ld{c.hitbox.as<uint64_t>(), cx, cy, cz};
      PartStore st{c.part_lists.as<uint32_t>() + c.part_used, d_tot};
      SVO_TRY((scan_exclusive<uint32_t>(ld, st, n, c.scan_scratch.as<uint32_t>(), c.stream)));
    }
    ull tot = 0;
    const ReadReq rq{&tot, d_tot, 8};
    SVO_TRY(read_back(c, &rq, 1));
    J.part_off[j] = c.part_used;
    J.part_cnt[j] = tot;
    c.part_used += tot;
  }
  return hipSuccess;
}

// The unit's brick-table segment covers its Morton range in bricks.
static uint64_t unit_bricks(const VoxUnit& u) {
  return (u.mend - (u.mstart & ~(uint64_t)(kBrickVox - 1)) + kBrickVox - 1) >> kBrickBits;
}

static void fill_params(const VoxUnit& h, const Context& c, VoxParams& P) {
  P.tris = h.tris;
  P.list = h.list;
  P.n = (uint32_t)h.n;
  P.stride = h.stride;
  P.mstart = h.mstart;
  P.bstart = h.mstart & ~(uint64_t)(kBrickVox - 1);
  uint32_t x, y, z;
  unmorton3(h.mstart, x, y, z);
  P.pmin[0] = (int)x;
  P.pmin[1] = (int)y;
  P.pmin[2] = (int)z;
  unmorton3(h.mend - 1, x, y, z);
  P.pmax[0] = (int)x;
  P.pmax[1] = (int)y;
  P.pmax[2] = (int)z;
  P.unit = h.unit;
  P.unit_div = 1.0f / h.unit;  // voxelizer.cpp:164, IEEE on the host as on the reference
  P.coord_limit = 4.0f * (float)h.gridsize * h.unit;
  // Along a column the voxels c that pass the reference's plane test lie, in exact
  // arithmetic, in [l, h] (the two zeros of (t + d1) and (t + d2)). Its FP32 evaluation
  // moves them by at most: t = n.p with |p| <= G u: 3 roundings of terms <= |n|_1 G u,
  // ~3.1e-7 G u; d1, d2 with |v0| <= coord_limit = 4 G u: ~1.7e-6 G u; so ~2e-6 G u, i.e.
  // 3.5e-6 G voxels along the column axis (|n_w| >= 1/sqrt(3)); p = fl(x u) adds ~3e-7 G.
  // This range is computed from float coefficients (KA, KB, alpha rounded to float, base
  // summed in float): below 1.8e-6 G voxels. The margin, 0.03 + 2.5e-5 G voxels, is over
  // four times the sum; the candidates are the integers in [l - margin, h + margin].
  P.slab_margin = 0.03 + 2.5e-5 * (double)h.gridsize;
  P.inv_unit = 1.0 / (double)h.unit;
  P.small_cols = c.tune.small_cols;
  P.direct_cols = c.tune.direct_cols;
  P.job = h.job;
  P.dense = 0;
  P.wide = 0;
  P.table_base = 0;
  P.grid8 = nullptr;
}

hipError_t voxelize_units(Context& c, const std::vector<VoxUnit>& hu, bool batch,
                          uint64_t* nfilled) {
  *nfilled = 0;
  std::vector<VoxUnit> units;
  for (const VoxUnit& u : hu)
    if (u.n) units.push_back(u);  // main.cpp:330 skips empty partitions
  const uint32_t nu = (uint32_t)units.size();
  if (nu == 0) return hipSuccess;
  // unit table: VoxParams[nu] | ubase[nu+1] | tile_off[nu+1] | uvoff[nu+1]
  std::vector<VoxParams> P(nu);
  std::vector<uint64_t> ubase(nu + 1);
  std::vector<uint32_t> toff(nu + 1);
  uint64_t ntab = 0, ntiles = 0, ntri = 0;
  // triangles per tile: kTileTris once the pass has about four tiles of them per block of
  // the persistent grid, else fewer (a multiple of 32, at least 32)
  // (passes with payload units: kTileTrisCompact, the tile of the compact item layout)
  uint32_t tt = (uint32_t)kTileTris;
  int small_cols = c.tune.small_cols;
  {
    uint64_t all = 0;
    bool any_payload_unit = false;
    for (uint32_t u = 0; u < nu; u++) {
      all += units[u].n;
      any_payload_unit |= units[u].stride == 21;
    }
    const uint64_t cap = any_payload_unit ? kTileTrisCompact : kTileTris;
    const uint64_t blocks = (uint64_t)std::max(1, c.tune.tile_blocks);
    const uint64_t want = all / (4 * blocks);
    tt = (uint32_t)std::min<uint64_t>(cap, std::max<uint64_t>(32, (want + 31) / 32 * 32));
    // A pass of one mesh with 16 or more tiles per block of the persistent grid has the
    // parallelism to walk triangles of up to 16 x small_cols columns in the small tier (rows
    // narrowed to their intervals, a triangle's segments shared by its block's threads),
    // which is faster than the large tier's thread per bbox column (the soup of 5x10^7
    // triangles at 8192, 8 partitions: 68.6 -> 61.1 ms per step). Batches keep their
    // autotuned threshold.
    if (!batch && all >= 16 * blocks * tt)
      small_cols = std::max(small_cols, (int)std::min<int64_t>(16ll * small_cols, 1 << 16));
  }
  for (uint32_t u = 0; u < nu; u++) {
    // a unit smaller than a brick shares its brick with its neighbours: the brick stream
    // would hold that brick more than once
    const uint64_t G = units[u].gridsize;
    if (units[u].mend - units[u].mstart < std::min<uint64_t>(kBrickVox, G * G * G))
      c.bricks_ok = false;
    if (units[u].n > 0xFFFFFFF0ull) {
      set_error("partition holds more than 2^32 triangles");
      return hipErrorInvalidValue;
    }
    fill_params(units[u], c, P[u]);
    P[u].small_cols = small_cols;
    P[u].table_base = ntab;
    ubase[u] = ntab;
    toff[u] = (uint32_t)ntiles;
    ntab += unit_bricks(units[u]);
    ntiles += (units[u].n + tt - 1) / tt;
    ntri += units[u].n;
  }
  ubase[nu] = ntab;
  toff[nu] = (uint32_t)ntiles;
  if (ntab > 0xFFFFFFF0ull || ntiles > 0xFFFFFFF0ull || ntri > 0xFFFFFFF0ull) {
    set_error("voxelize pass too large for 32-bit brick, tile or triangle indices");
    return hipErrorInvalidValue;
  }
  // dense path: a cell per voxel of every unit (a byte, 4 bytes for payload units) when
  // that fits the HBM budget (no allocation pass needed); otherwise bricks get pool slots
  // from a touch pass
  std::vector<uint64_t> goff(nu + 1, 0);
  for (uint32_t u = 0; u < nu; u++)
    goff[u + 1] = goff[u] + ((ubase[u + 1] - ubase[u]) << kBrickBits) *
                                (units[u].stride == 21 ? 4 : 1);
  const uint64_t dense_bytes = goff[nu];
  const bool dense = dense_bytes <= c.dense_limit;
  bool geometry_only = true;  // no unit carries payload records
  for (uint32_t u = 0; u < nu; u++) geometry_only = geometry_only && units[u].stride == 9;
  // the table, the dense grid and the byte pool are all-zero whenever no pass is running
  if (c.table_entries < ntab) {
    SVO_TRY(ensure(c.table, ntab * 4));
    SVO_TRY(fill_u32(c.table.as<uint32_t>(), 0, ntab, c.stream));
    c.table_entries = ntab;
  }
  if (dense) {
    if (c.grid_zeroed < dense_bytes) {
      // grown to at least twice its size (at most the dense limit), so that passes of
      // slowly growing spans (a server's batches) reallocate it a few times only
      uint64_t target = dense_bytes;
      if (c.grid8.cap < dense_bytes) {
        target = std::max<uint64_t>(dense_bytes,
                                    std::min<uint64_t>(c.dense_limit, 2 * (uint64_t)c.grid8.cap));
        release(c.grid8);
        SVO_TRY(ensure(c.grid8, target));
      }
      SVO_TRY(fill_u64(c.grid8.as<uint64_t>(), 0, target / 8, c.stream));
      c.grid_zeroed = target;
    }
    for (uint32_t u = 0; u < nu; u++) {
      P[u].dense = 1;
      P[u].wide = units[u].stride == 21 ? 1 : 0;
      P[u].grid8 = c.grid8.as<uint8_t>() + goff[u];
    }
  } else if (c.pool_cap == 0) {
    const uint64_t cap = std::max<uint64_t>(1024, ntri / 64);
    SVO_TRY(ensure(c.pool, cap * kBrickWords * 8));
    SVO_TRY(ensure(c.pool8, cap * kBrickVox));
    SVO_TRY(fill_u64(c.pool8.as<uint64_t>(), 0, cap * kBrickVox / 8, c.stream));
    c.pool_cap = cap;
  }
  const size_t szP = ((nu * sizeof(VoxParams)) + 255) & ~(size_t)255;
  const size_t szB = (((nu + 1) * 8) + 255) & ~(size_t)255;
  const size_t szT = (((nu + 1) * 4) + 255) & ~(size_t)255;
  {
    const void* prev = c.unitdev.p;
    const size_t prev_cap = c.unitdev.cap;
    SVO_TRY(ensure(c.unitdev, szP + 2 * szB + szT));
    if (c.unitdev.p != prev || c.unitdev.cap != prev_cap) c.unit_img.clear();
  }
  VoxParams* d_units = (VoxParams*)c.unitdev.p;
  uint64_t* d_ubase = (uint64_t*)((char*)c.unitdev.p + szP);
  uint32_t* d_toff = (uint32_t*)((char*)c.unitdev.p + szP + szB);
  uint64_t* d_uvoff = (uint64_t*)((char*)c.unitdev.p + szP + szB + szT);
  {  // one upload for the unit table (same layout as on the device)
    std::vector<uint8_t> img(szP + szB + szT, 0);
    memcpy(img.data(), P.data(), nu * sizeof(VoxParams));
    memcpy(img.data() + szP, ubase.data(), (nu + 1) * 8);
    memcpy(img.data() + szP + szB, toff.data(), (nu + 1) * 4);
    // a pass over the same units as the last one (the same mesh again) finds them there
    if (img != c.unit_img) {
      SVO_TRY(upload_async(c, d_units, img.data(), img.size()));
      c.unit_img.swap(img);
    }
  }

  // dense passes write their bricks straight into the brick stream (every flagged brick
  // holds voxels): room for all of the pass's bricks after the nb0 already there
  const uint64_t s0 = c.n_vox;
  const uint64_t nb0 = c.n_bricks;
  if (dense) {
    SVO_TRY(ensure_keep(c.bwords, (nb0 + ntab) * kBrickWords * 8, nb0 * kBrickWords * 8,
                        c.stream));
    SVO_TRY(ensure_keep(c.bcode, (nb0 + ntab) * 8, nb0 * 8, c.stream));
    SVO_TRY(ensure_keep(c.bvoff, (nb0 + ntab) * 8, nb0 * 8, c.stream));
    SVO_TRY(ensure_keep(c.bshape, (nb0 + ntab) * 4, nb0 * 4, c.stream));
  }
  ull* pool_words = dense ? c.bwords.as<ull>() + nb0 * kBrickWords : c.pool.as<ull>();
  SVO_TRY(ensure(c.large_unit, ntri * 4));
  SVO_TRY(ensure(c.large_tri, ntri * 4));
  SVO_TRY(ensure(c.large_cols, ntri * 4));
  SVO_TRY(ensure(c.large_off, ntri * 8 + 8));
  // walk constants of the first kLargeWalkMax large triangles of a pass (the others are
  // set up again by each of their column threads)
  const uint32_t walk_cap = (uint32_t)std::min<uint64_t>(ntri, kLargeWalkMax);
  SVO_TRY(ensure(c.large_walk, (size_t)walk_cap * kWalkWords * 4 + 4));
  SVO_TRY(ensure(c.occ_list, ntab * 4));
  SVO_TRY(ensure_scan_scratch(c, std::max<uint64_t>(ntri, ntab), 8));
  ull* ctr = c.counters.as<ull>();
  ull hc[kNumCounters];
  const unsigned tile_grid = (unsigned)std::min<uint64_t>(ntiles, c.tune.tile_blocks);
  const unsigned large_grid = (unsigned)c.tune.large_blocks;
  auto brick_grid = [&](uint64_t bricks) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((bricks * kWave + 255) / 256, 8192));
  };
  auto args = [&]() {
    ActArgs A;
    A.table = c.table.as<uint32_t>();
    A.pool8 = c.pool8.as<uint8_t>();
    A.pool = pool_words;
    A.wprefix = c.wprefix.as<uint32_t>();
    A.winner = c.winner.as<uint32_t>();
    A.ctr = ctr;
    A.pool_cap = c.pool_cap;
    return A;
  };
  uint32_t* l_unit = c.large_unit.as<uint32_t>();
  uint32_t* l_tri = c.large_tri.as<uint32_t>();
  uint32_t* l_cols = c.large_cols.as<uint32_t>();
  uint64_t* l_off = c.large_off.as<uint64_t>();
  uint32_t* l_walk = c.large_walk.as<uint32_t>();
  auto tile_pass = [&](auto tag, const char* name, ull* queue, int collect, int skip) {
    using Act = decltype(tag);
    ktimer_begin(c, name);
    SVO_LAUNCH((k_vox_tile<Act>), dim3(tile_grid), dim3(kTile), 0, c.stream, d_units, nu,
                       d_toff, args(), queue, l_unit, l_tri, l_cols, l_walk, walk_cap, tt,
                       collect, skip);
    ktimer_end(c);
    return hipGetLastError();
  };
  auto large_pass = [&](auto tag, const char* name, int skip) {
    using Act = decltype(tag);
    ktimer_begin(c, name);
    SVO_LAUNCH((k_vox_large<Act>), dim3(large_grid), dim3(256), 0, c.stream, d_units,
                       args(), l_unit, l_tri, l_off, l_walk, walk_cap, skip);
    ktimer_end(c);
    return hipGetLastError();
  };
  // single-pass scans, each with its own zeroed status words: large-triangle columns,
  // flagged bricks, brick counts (the last only while its packed prefixes stay below 2^62:
  // at most bricks * (2^kNonEmptyShift + 4096))
  const uint64_t w_col = scan_single_words(ntri), w_occ = scan_single_words(ntab);
  ull* st_col = nullptr;
  ull* st_occ = nullptr;
  ull* st_bc = nullptr;
  ull* queues = nullptr;  // 3 x kQueueWords: the tile queues of the mark/touch, second and winner passes
  auto occ_scan = [&]() {
    return scan_single<uint32_t>(
        OccLoad{c.table.as<uint32_t>()},
        OccStore{c.occ_list.as<uint32_t>(), c.table.as<uint32_t>(), ctr}, ntab, st_occ,
        ctr + kCtrError, c.stream);
  };

  // ---- phase A: mark, compact counts; one synchronization at its end
  std::vector<uint64_t> uvoff(nu + 1);
  for (int attempt = 0;; attempt++) {
    const uint64_t bc_bound = dense ? ntab : std::min<uint64_t>(ntab, c.pool_cap);
    const bool bc_single = bc_bound <= (3ull << 20);
    const uint64_t w_bc = bc_single ? scan_single_words(bc_bound) : 0;
    SVO_TRY(ensure(c.bcount, bc_bound * 8 + 8));
    // status words of the scans, then the tile queues of the three tile passes
    SVO_TRY(ensure(c.scan_status, (w_col + w_occ + w_bc + 3 * kQueueWords) * 8));
    st_col = c.scan_status.as<ull>();
    st_occ = st_col + w_col;
    st_bc = st_occ + w_occ;
    queues = st_bc + w_bc;
    // the pass's kernels up to its read-back: one graph segment
    SVO_TRY(run_segment(c, [&]() -> hipError_t {
      {  // pass counters, scan status words and tile queues
        const uint64_t ns = w_col + w_occ + w_bc + 3 * kQueueWords;
        SVO_LAUNCH(k_zero2_u64, dim3(fill_grid(kCtrPassClear + ns)), dim3(256), 0,
                           c.stream, (uint64_t*)ctr, (uint64_t)kCtrPassClear,
                           c.scan_status.as<uint64_t>(), ns);
        SVO_TRY(hipGetLastError());
      }
      // first pass: dense units mark directly, touch units flag their bricks; both collect
      // the large triangles
      // payload passes: the lane-neighbour check before the atomic pays on one mesh (a fine
      // payload terrain at 1024: mark 357 -> 213 us) but cost the 512-mesh batched mix 11%
      if (dense && geometry_only)
        SVO_TRY(tile_pass(MarkDenseAction{}, "vox_small_mark", queues, 1, 0));
      else if (dense && nu == 1)
        SVO_TRY(tile_pass(MarkDenseWideAction{}, "vox_small_mark", queues, 1, 0));
      else if (dense)
        SVO_TRY(tile_pass(MarkDenseWideBatchAction{}, "vox_small_mark", queues, 1, 0));
      else
        SVO_TRY(tile_pass(TouchAction{}, "vox_small_touch", queues, 1, 0));
      SVO_TRY((scan_single<uint64_t>(ColLoad{l_cols},
                                     CountStore{l_off, ctr + kCtrCols},
                                     ntri, st_col, ctr + kCtrError, c.stream,
                                     ctr + kCtrLargeCount)));
      if (dense) {
        if (geometry_only)
          SVO_TRY(large_pass(MarkDenseAction{}, "vox_large_mark", 0));
        else  // (the large tier's lanes are not in triangle order)
          SVO_TRY(large_pass(MarkDenseWideBatchAction{}, "vox_large_mark", 0));
        // number the flagged bricks in Morton order: table entry = rank + 1, occ = list
        SVO_TRY(occ_scan());
      } else {
        SVO_TRY(large_pass(TouchAction{}, "vox_large_touch", 0));
        SVO_TRY(occ_scan());
        // mark pass: byte stores into the slots
        SVO_TRY(tile_pass(MarkPoolAction{}, "vox_small_mark", queues + kQueueWords, 0, 0));
        SVO_TRY(large_pass(MarkPoolAction{}, "vox_large_mark", 0));
      }
      // per-brick voxel counts (bytes packed to bits, bytes cleared) and their offsets
      ktimer_begin(c, "brick_count");
      SVO_LAUNCH(k_brick_count, dim3(brick_grid(bc_bound)), dim3(256), 0, c.stream,
                         c.occ_list.as<uint32_t>(), ctr, bc_bound, d_units, d_ubase, nu,
                         c.pool8.as<uint8_t>(), pool_words, c.bcount.as<uint64_t>(),
                         dense ? c.bcode.as<uint64_t>() + nb0 : (uint64_t*)nullptr,
                         dense ? c.bshape.as<uint32_t>() + nb0 : (uint32_t*)nullptr,
                         dense ? c.table.as<uint32_t>() : (uint32_t*)nullptr);
      ktimer_end(c);
      SVO_TRY(hipGetLastError());
      if (bc_single && dense)
        SVO_TRY((scan_single<uint64_t>(
            ArrayLoad<uint64_t>{c.bcount.as<uint64_t>()},
            BrickOffStore{c.bcount.as<uint64_t>(), ctr + kCtrNumVox,
                          c.bvoff.as<uint64_t>() + nb0, s0},
            bc_bound, st_bc, ctr + kCtrError, c.stream, ctr + kCtrNumOcc)));
      else if (bc_single)
        SVO_TRY((scan_single<uint64_t>(
            ArrayLoad<uint64_t>{c.bcount.as<uint64_t>()},
            CountStore{c.bcount.as<uint64_t>(), ctr + kCtrNumVox}, bc_bound,
            st_bc, ctr + kCtrError, c.stream, ctr + kCtrNumOcc)));
      else if (dense)
        SVO_TRY((scan_exclusive<uint64_t>(
            ArrayLoad<uint64_t>{c.bcount.as<uint64_t>()},
            BrickOffStore{c.bcount.as<uint64_t>(), ctr + kCtrNumVox,
                          c.bvoff.as<uint64_t>() + nb0, s0},
            bc_bound, c.scan_scratch.as<uint64_t>(), c.stream, ctr + kCtrNumOcc)));
      else
        SVO_TRY((scan_exclusive<uint64_t>(
            ArrayLoad<uint64_t>{c.bcount.as<uint64_t>()},
            CountStore{c.bcount.as<uint64_t>(), ctr + kCtrNumVox}, bc_bound,
            c.scan_scratch.as<uint64_t>(), c.stream, ctr + kCtrNumOcc)));
      if (!(nu == 1 && geometry_only)) {
        SVO_LAUNCH(k_unit_voxels, dim3((nu + 256) / 256), dim3(256), 0, c.stream,
                   c.occ_list.as<uint32_t>(), ctr, d_ubase, nu, c.bcount.as<uint64_t>(),
                   d_uvoff);
        SVO_TRY(hipGetLastError());
      }
      return hipSuccess;
    }));
    if (nu == 1 && geometry_only) {
      // one unit and no payload records (the only device reader of the unit offsets):
      // its voxels are the whole pass, [0, voxel count)
      const ReadReq rq = {hc, c.counters.p, kNumCounters * sizeof(ull)};
      SVO_TRY(read_back(c, &rq, 1));
      uvoff[0] = 0;
      uvoff[1] = hc[kCtrNumVox] & kVoxMask;
    } else {
      const ReadReq rq[2] = {{uvoff.data(), d_uvoff, (nu + 1) * 8},
                             {hc, c.counters.p, kNumCounters * sizeof(ull)}};
      SVO_TRY(read_back(c, rq, 2));
    }
    ktimer_flush(c);
    if (dense || (!hc[kCtrOverflow] && hc[kCtrNumOcc] <= c.pool_cap)) break;
    // pool too small for the flagged bricks: clear everything, grow, redo
    const uint64_t want = std::max<uint64_t>(c.pool_cap * 2, hc[kCtrNumOcc] + 1024);
    if (attempt > 8 || want > (1ull << 26)) {
      set_error("brick pool could not be sized");
      return hipErrorOutOfMemory;
    }
    SVO_TRY(fill_u32(c.table.as<uint32_t>(), 0, ntab, c.stream));
    release(c.pool);
    release(c.pool8);
    SVO_TRY(ensure(c.pool, want * kBrickWords * 8));
    SVO_TRY(ensure(c.pool8, want * kBrickVox));
    SVO_TRY(fill_u64(c.pool8.as<uint64_t>(), 0, want * kBrickVox / 8, c.stream));
    c.pool_cap = want;
  }
  const uint64_t n_occ = hc[kCtrNumOcc];
  const uint64_t nvox = hc[kCtrNumVox] & kVoxMask;
  const uint64_t n_nonempty = hc[kCtrNumVox] >> kNonEmptyShift;
  const uint64_t n_large = hc[kCtrLargeCount];
  if (c.n_vox + nvox >= (1ull << 40)) {
    set_error("voxel stream too long");
    return hipErrorInvalidValue;
  }

  // ---- job layout: voxel and data-record ranges
  const bool any_payload = !geometry_only;
  if (batch) {
    // the passes of a batch take its jobs' partitions in job order: a job's range starts
    // at its first voxel and grows with every pass that holds its partitions (jobs start
    // with voff = ~0, nvox = 0; the caller places jobs without voxels afterwards)
    for (uint32_t u = 0; u < nu; u++) {
      Job& J = c.jobs[units[u].job];
      const uint64_t k = uvoff[u + 1] - uvoff[u];
      if (J.voff == ~0ull && k) J.voff = s0 + uvoff[u];
      J.nvox += k;
    }
  } else {
    Job& J = c.jobs[0];
    J.voff = 0;
    J.nvox = s0 + nvox;
  }
  // data ranges: a job's records start after the complete ranges of the jobs before it
  // (jobs after the last one of this pass have no records yet, their bases move later)
  uint64_t n_records = 0;
  for (size_t j = 0; j < c.jobs.size(); j++) {
    Job& J = c.jobs[j];
    J.dbase = n_records;
    J.n_data = J.payload ? 1 + J.nvox : 2;
    n_records += J.n_data;
  }
  std::vector<JobDev> jd(c.jobs.size());
  for (size_t j = 0; j < c.jobs.size(); j++) {
    const Job& J = c.jobs[j];
    JobDev& d = jd[j];
    memset(&d, 0, sizeof(d));
    d.tris = J.tris;
    d.unit_div = 1.0f / J.unit_vox;
    d.color_mode = J.color_mode;
    d.payload = J.payload;
    d.gridsize = J.gridsize;
    int D = 0;
    while ((1ull << D) < J.gridsize) D++;
    d.depth = D;
    d.voff = J.voff == ~0ull ? 0 : J.voff;
    d.nvox = J.nvox;
    d.dbase = J.dbase;
  }
  SVO_TRY(ensure(c.jobdev, jd.size() * sizeof(JobDev)));
  SVO_TRY(upload_async(c, c.jobdev.p, jd.data(), jd.size() * sizeof(JobDev)));
  c.jobdev_img = jd;

  // ---- phase B: emit sorted codes, payload, clear; no synchronization
  if (n_occ) {
    // dense passes resolve payload winners while marking (emitted with the codes); the
    // touched-brick path finds them with a second walk
    const bool winner_walk = any_payload && !dense;
    if (any_payload && nvox > 0xFFFFFFF0ull) {
      set_error("voxelize pass holds more than 2^32 voxels");
      return hipErrorInvalidValue;
    }
    if (winner_walk) SVO_TRY(ensure(c.wprefix, n_occ * kBrickWords * 4));
    if (any_payload) SVO_TRY(ensure(c.winner, nvox * 4));
    // geometry-only passes write no codes: the build reads the brick stream, and
    // ensure_codes writes the codes for readers that want them
    const bool codes_now = any_payload;
    SVO_TRY(ensure_keep(c.codes, (s0 + nvox) * 8, s0 * 8, c.stream));
    if (!dense) {
      SVO_TRY(ensure_keep(c.bwords, (nb0 + n_occ) * kBrickWords * 8, nb0 * kBrickWords * 8,
                          c.stream));
      SVO_TRY(ensure_keep(c.bcode, (nb0 + n_occ) * 8, nb0 * 8, c.stream));
      SVO_TRY(ensure_keep(c.bvoff, (nb0 + n_occ) * 8, nb0 * 8, c.stream));
      SVO_TRY(ensure_keep(c.bshape, (nb0 + n_occ) * 4, nb0 * 4, c.stream));
    }
    // the payload units' voxel ranges, for the records kernel: prange (n_pu + 1) |
    // pstart (n_pu) | punit (n_pu, as u64)
    uint64_t n_pv = 0;
    uint32_t n_pu = 0;
    const uint64_t* d_prange = nullptr;
    const uint64_t* d_pstart = nullptr;
    const uint32_t* d_punit = nullptr;
    if (any_payload) {
      // records: each job's first record is its null record
      const uint64_t keep = batch ? c.records.cap : (1 + s0) * 32;
      SVO_TRY(ensure_keep(c.records, n_records * 32, keep, c.stream));
      std::vector<uint64_t> pr(1, 0), ps, pu;
      for (uint32_t u = 0; u < nu; u++) {
        const uint64_t cnt = uvoff[u + 1] - uvoff[u];
        if (units[u].stride != 21 || cnt == 0) continue;
        pr.push_back(pr.back() + cnt);
        ps.push_back(uvoff[u]);
        pu.push_back(u);
      }
      n_pu = (uint32_t)ps.size();
      n_pv = pr.back();
      if (n_pv) {
        std::vector<uint64_t> img(pr);
        img.insert(img.end(), ps.begin(), ps.end());
        std::vector<uint32_t> pu32(pu.begin(), pu.end());
        const size_t b64 = img.size() * 8;
        SVO_TRY(ensure(c.payload_map, b64 + pu32.size() * 4));
        SVO_TRY(upload_async(c, c.payload_map.p, img.data(), b64));
        SVO_TRY(upload_async(c, (char*)c.payload_map.p + b64, pu32.data(), pu32.size() * 4));
        d_prange = c.payload_map.as<uint64_t>();
        d_pstart = d_prange + n_pu + 1;
        d_punit = (const uint32_t*)((const char*)c.payload_map.p + b64);
      }
    }
    // the pass's remaining kernels: one graph segment
    SVO_TRY(run_segment(c, [&]() -> hipError_t {
      // a dense pass has its brick stream entries already (k_brick_count); the emit pass
      // writes the codes and winners of its payload units
      if (!dense || codes_now) {
        ktimer_begin(c, "brick_emit");
        // dense payload units: their cells are staged through LDS, kBrickVox words per wave
        const bool stage_cells = codes_now && dense && any_payload;
        SVO_LAUNCH(codes_now ? k_brick_emit<true> : k_brick_emit<false>,
                   dim3(brick_grid(n_occ)), dim3(256),
                   stage_cells ? (256 / kWave) * kBrickVox * 4 : 0, c.stream,
                   c.occ_list.as<uint32_t>(), ctr, d_units, d_ubase, nu, pool_words,
                   c.bcount.as<uint64_t>(), c.codes.as<uint64_t>() + s0,
                   winner_walk ? c.wprefix.as<uint32_t>() : (uint32_t*)nullptr, s0,
                   c.bwords.as<uint64_t>() + nb0 * kBrickWords, c.bcode.as<uint64_t>() + nb0,
                   c.bvoff.as<uint64_t>() + nb0, c.bshape.as<uint32_t>() + nb0,
                   any_payload ? c.winner.as<uint32_t>() : (uint32_t*)nullptr,
                   stage_cells ? 1 : 0, dense ? 0 : 1);
        ktimer_end(c);
        SVO_TRY(hipGetLastError());
      }
      if (any_payload) {
        if (winner_walk) {
          SVO_TRY(fill_u32(c.winner.as<uint32_t>(), 0xFFFFFFFFu, nvox, c.stream));
          SVO_TRY(tile_pass(WinnerAction{}, "vox_small_winner", queues + 2 * kQueueWords, 0,
                            kSkipBinary));
          if (n_large) SVO_TRY(large_pass(WinnerAction{}, "vox_large_winner", kSkipBinary));
        }
        if (n_pv) {
          ktimer_begin(c, "payload_records");
          SVO_LAUNCH(k_payload_records, dim3((unsigned)((n_pv + 255) / 256)), dim3(256), 0,
                     c.stream, c.jobdev.as<JobDev>(), (const VoxParams*)d_units, d_prange,
                     d_pstart, d_punit, n_pu, c.codes.as<uint64_t>(), s0,
                     c.winner.as<uint32_t>(), n_pv, c.records.as<float>(), ctr);
          ktimer_end(c);
          SVO_TRY(hipGetLastError());
        }
      }
      if (!dense) {  // (dense passes cleared their table entries in k_brick_count)
        ktimer_begin(c, "brick_clear");
        SVO_LAUNCH(k_brick_clear,
                   dim3((unsigned)std::min<uint64_t>((n_occ + 255) / 256, 4096)),
                   dim3(256), 0, c.stream, c.occ_list.as<uint32_t>(), ctr,
                   c.table.as<uint32_t>());
        ktimer_end(c);
        SVO_TRY(hipGetLastError());
      }
      return hipSuccess;
    }));
    c.n_bricks = nb0 + n_nonempty;
    bool any_binary = false;
    for (uint32_t u = 0; u < nu; u++) any_binary |= units[u].stride == 9;
    if (!codes_now || any_binary) c.codes_lazy = true;
  }
  c.n_vox = s0 + nvox;
  *nfilled = nvox;
  return hipSuccess;
}

hipError_t ensure_codes(Context& c) {
  if (!c.codes_lazy) return hipSuccess;
  // every brick again (rewriting codes already there with the same values)
  if (c.n_bricks) {
    const unsigned grid =
        (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((c.n_bricks * kWave + 255) / 256, 8192));
    SVO_LAUNCH(k_bricks_codes, dim3(grid), dim3(256), 0, c.stream,
                       c.bwords.as<uint64_t>(), c.bcode.as<uint64_t>(), c.bvoff.as<uint64_t>(),
                       c.n_bricks, c.codes.as<uint64_t>());
    SVO_TRY(hipGetLastError());
  }
  c.codes_lazy = false;
  return hipSuccess;
}

}  // namespace svo
