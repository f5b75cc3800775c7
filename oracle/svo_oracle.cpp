// ORACLE — test infrastructure, not product code.
//
// A from-scratch, single-threaded CPU restatement of ooc_svo_builder v1.6.4's
// partition -> voxelize -> SVO-build path. Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load it: as the checker the GPU library is compared
// against, and as the timed CPU baseline. Nothing in the product links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to the
// reference repository root). FP32 arithmetic follows the reference's expression
// order, with glm's vector routines restated as in glm 0.9.9:
//   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)    = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
//   normalize(v)  = v * (1.0f / sqrt(dot(v,v)))
//   inverse(mat3) = cofactors * (1 / det), det expanded along column 0
//   mat3 * vec3   = row r: (m[0][r]*v.x + m[1][r]*v.y) + m[2][r]*v.z
// glm itself is absent from the image, so agreement with the real glm build is
// "parity unpinned" for the FP32 parts (normal, plane/edge constants, barycentric
// colours); the integer parts (Morton, partition counts, octree structure and byte
// layout) are pinned by closed forms in tests/.
//
// Build: oracle/Makefile, g++ -O3 -ffp-contract=off (no -march, like the reference's
// linux/build_svo_builder.sh:9, so no FMA); x86-64 SSE, FLT_EVAL_METHOD 0.
//
// Deliberate, documented departures from the reference (undefined behaviour there):
//  * finalizeTree with the cursor stopped exactly at the last Morton code
//    (OctreeBuilder.cpp:36-42 reads an empty buffer) and an empty 1^3 grid: the tree
//    is completed with empty padding and hazard() reports it.
//  * NaN values written to .octreedata are stored as 0xFFC00000, the x86 default NaN.
//  * a .tridata shorter than its header claims is an error, not garbage triangles.
#include "svo_oracle.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <memory>
#include <vector>

namespace orc {

// ------------------------------------------------------------------------------------
// glm-like FP32 vectors (restating the glm routines the reference calls)
// ------------------------------------------------------------------------------------
struct vec3 {
  float x, y, z;
  float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
struct vec2 {
  float x, y;
};
static inline vec3 mk3(float a, float b, float c) { return vec3{a, b, c}; }
static inline vec3 operator-(vec3 a, vec3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline vec3 operator+(vec3 a, vec3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline vec3 operator*(float s, vec3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
static inline vec3 operator*(vec3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline vec3 operator/(vec3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline vec2 operator*(float s, vec2 a) { return vec2{s * a.x, s * a.y}; }
static inline float dot(vec3 a, vec3 b) {
  float tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
  return (tx + ty) + tz;
}
static inline float dot(vec2 a, vec2 b) {
  float tx = a.x * b.x, ty = a.y * b.y;
  return tx + ty;
}
static inline vec3 cross(vec3 a, vec3 b) {
  return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline vec3 normalize(vec3 v) {
  float inv = 1.0f / std::sqrt(dot(v, v));
  return v * inv;
}
// std::min / std::max on floats: min(a,b) = (b<a)?b:a, max(a,b) = (a<b)?b:a
static inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_std(float a, float b) { return (a < b) ? b : a; }

struct mat3 {
  vec3 c[3];  // columns; m[c][r]
};
static inline mat3 inverse(const mat3& M) {
  const float m00 = M.c[0].x, m01 = M.c[0].y, m02 = M.c[0].z;
  const float m10 = M.c[1].x, m11 = M.c[1].y, m12 = M.c[1].z;
  const float m20 = M.c[2].x, m21 = M.c[2].y, m22 = M.c[2].z;
  float det = (m00 * (m11 * m22 - m21 * m12) - m10 * (m01 * m22 - m21 * m02)) +
              m20 * (m01 * m12 - m11 * m02);
  float o = 1.0f / det;
  mat3 I;
  I.c[0].x = (m11 * m22 - m21 * m12) * o;
  I.c[1].x = (-(m10 * m22 - m20 * m12)) * o;
  I.c[2].x = (m10 * m21 - m20 * m11) * o;
  I.c[0].y = (-(m01 * m22 - m21 * m02)) * o;
  I.c[1].y = (m00 * m22 - m20 * m02) * o;
  I.c[2].y = (-(m00 * m21 - m20 * m01)) * o;
  I.c[0].z = (m01 * m12 - m11 * m02) * o;
  I.c[1].z = (-(m00 * m12 - m10 * m02)) * o;
  I.c[2].z = (m00 * m11 - m10 * m01) * o;
  return I;
}
static inline vec3 mul(const mat3& m, vec3 v) {
  return mk3((m.c[0].x * v.x + m.c[1].x * v.y) + m.c[2].x * v.z,
             (m.c[0].y * v.x + m.c[1].y * v.y) + m.c[2].y * v.z,
             (m.c[0].z * v.x + m.c[1].z * v.y) + m.c[2].z * v.z);
}

// static_cast<int>(float) as x86 cvttss2si executes it: NaN and out-of-range give INT_MIN
static inline int trunc_i32_x86(float f) {
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT_MIN;
  return (int)f;
}

static inline uint32_t f2u(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
static inline float u2f(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static inline float canon_nan(float f) { return (f != f) ? u2f(0xFFC00000u) : f; }

// ------------------------------------------------------------------------------------
// Morton codes: libmorton/include/morton3D.h:36-47 (encode, shifted byte LUTs, 8 rounds
// of <<24 because the coordinate type is 8 bytes on LP64) and :163-178 (decode, 7
// rounds of 9-bit LUTs). The tables (morton3D_LUTs.h:7-224) are rebuilt here from
// their closed form: byte bit i -> code bit 3i (+1 for y, +2 for z).
// ------------------------------------------------------------------------------------
struct MortonTables {
  uint64_t enc_x[256], enc_y[256], enc_z[256];
  uint8_t dec_x[512], dec_y[512], dec_z[512];
  MortonTables() {
    for (uint32_t b = 0; b < 256; b++) {
      uint64_t s = 0;
      for (int i = 0; i < 8; i++)
        if (b & (1u << i)) s |= 1ull << (3 * i);
      enc_x[b] = s;
      enc_y[b] = s << 1;
      enc_z[b] = s << 2;
    }
    for (uint32_t c = 0; c < 512; c++) {
      uint8_t x = 0, y = 0, z = 0;
      for (int i = 0; i < 3; i++) {
        if (c & (1u << (3 * i))) x |= (uint8_t)(1u << i);
        if (c & (1u << (3 * i + 1))) y |= (uint8_t)(1u << i);
        if (c & (1u << (3 * i + 2))) z |= (uint8_t)(1u << i);
      }
      dec_x[c] = x;
      dec_y[c] = y;
      dec_z[c] = z;
    }
  }
};
static const MortonTables& tables() {
  static MortonTables t;
  return t;
}

static inline uint64_t morton_encode(uint64_t x, uint64_t y, uint64_t z) {
  const MortonTables& T = tables();
  uint64_t a = 0;
  for (unsigned i = 8; i > 0; --i) {
    unsigned s = (i - 1) * 8;
    a = (a << 24) | (T.enc_z[(z >> s) & 0xFF] | T.enc_y[(y >> s) & 0xFF] | T.enc_x[(x >> s) & 0xFF]);
  }
  return a;
}
static inline uint64_t decode_axis(uint64_t m, const uint8_t* lut) {
  uint64_t a = 0;
  for (unsigned i = 0; i < 7; ++i) a |= (uint64_t)(lut[(m >> (i * 9)) & 0x1FF] << (3 * i));
  return a;
}
static inline void morton_decode(uint64_t m, uint64_t& x, uint64_t& y, uint64_t& z) {
  const MortonTables& T = tables();
  x = decode_axis(m, T.dec_x);
  y = decode_axis(m, T.dec_y);
  z = decode_axis(m, T.dec_z);
}

// ------------------------------------------------------------------------------------
// svo_builder_util.h:14-52
// ------------------------------------------------------------------------------------
static inline unsigned find_power_of_8(uint64_t n) {  // :28-35
  if (n == 0) return 0;
  unsigned h = 0;
  while (n >>= 1) h++;
  return h / 3;
}
static inline unsigned log2_u(unsigned v) {  // :37-44
  unsigned r = (unsigned)-1;
  while (v != 0) {
    v >>= 1;
    r++;
  }
  return r;
}
static inline bool is_pow2(unsigned x) { return x != 0 && !(x & (x - 1)); }  // :46-48
static inline int clampval(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline vec3 morton_to_rgb(uint64_t m, size_t G) {  // :14-18 (note the swapped axes)
  uint64_t a, b, c;
  morton_decode(m, a, b, c);  // reference names them z, y, x
  const float g = (float)G;
  return mk3((float)c / g, (float)b / g, (float)a / g);
}

// ------------------------------------------------------------------------------------
// Records: tri_util.h:7-11,29-54 (Triangle, 9 or 21 floats), VoxelData.h:10-18 (32 B),
// Node.h:6-65 (24 B on disk: data, children_base, offsets[8]).
// ------------------------------------------------------------------------------------
struct Tri {
  vec3 v0, v1, v2, normal, c0, c1, c2;
};
static inline Tri load_tri(const float* p, bool payload) {
  Tri t;
  t.v0 = mk3(p[0], p[1], p[2]);
  t.v1 = mk3(p[3], p[4], p[5]);
  t.v2 = mk3(p[6], p[7], p[8]);
  if (payload) {
    t.normal = mk3(p[9], p[10], p[11]);
    t.c0 = mk3(p[12], p[13], p[14]);
    t.c1 = mk3(p[15], p[16], p[17]);
    t.c2 = mk3(p[18], p[19], p[20]);
  } else {
    t.normal = t.c0 = t.c1 = t.c2 = mk3(0, 0, 0);
  }
  return t;
}

struct VoxelData {
  uint64_t morton;
  vec3 color;
  vec3 normal;
};
static_assert(sizeof(VoxelData) == 32, "VoxelData is 32 bytes");
static inline VoxelData vd_zero() { return VoxelData{0, mk3(0, 0, 0), mk3(0, 0, 0)}; }

struct Node {
  uint64_t data = 0;
  uint64_t children_base = 0;
  int8_t off[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  VoxelData cache = vd_zero();
  bool is_leaf() const {
    for (int k = 0; k < 8; k++)
      if (off[k] != -1) return false;
    return true;
  }
  bool is_null() const { return is_leaf() && data == 0; }
};

// ------------------------------------------------------------------------------------
// Output sink: octree_io.h:49-66 (writeVoxelData 32 B, writeNode 24 B, return the
// pre-increment index). Either in memory or to FILE*.
// ------------------------------------------------------------------------------------
struct Sink {
  std::vector<uint8_t> nodes, data;
  FILE* fn = nullptr;
  FILE* fd = nullptr;
  uint64_t node_pos = 0, data_pos = 0;
  uint64_t write_node(const Node& n) {
    uint8_t rec[24];
    std::memcpy(rec, &n.data, 8);
    std::memcpy(rec + 8, &n.children_base, 8);
    std::memcpy(rec + 16, n.off, 8);
    if (fn)
      std::fwrite(rec, 24, 1, fn);
    else
      nodes.insert(nodes.end(), rec, rec + 24);
    return node_pos++;
  }
  uint64_t write_data(const VoxelData& v) {
    uint8_t rec[32];
    float f[6] = {canon_nan(v.color.x),  canon_nan(v.color.y),  canon_nan(v.color.z),
                  canon_nan(v.normal.x), canon_nan(v.normal.y), canon_nan(v.normal.z)};
    std::memcpy(rec, &v.morton, 8);
    std::memcpy(rec + 8, f, 24);
    if (fd)
      std::fwrite(rec, 32, 1, fd);
    else
      data.insert(data.end(), rec, rec + 32);
    return data_pos++;
  }
};

// ------------------------------------------------------------------------------------
// OctreeBuilder: OctreeBuilder.cpp:4-168, OctreeBuilder.h:50-91. A streaming,
// bottom-up builder over a strictly increasing Morton stream.
// ------------------------------------------------------------------------------------
class Builder {
 public:
  Builder(Sink* sink, size_t gridlength, bool payload, bool levels)
      : sink_(sink), G_(gridlength), payload_(payload), levels_(levels) {
    D_ = (int)log2_u((unsigned)gridlength);  // OctreeBuilder.cpp:15
    buf_.resize(D_ + 1);
    for (auto& b : buf_) b.reserve(8);
    uint64_t m = gridlength - 1;
    max_ = morton_encode(m, m, m);  // :22-23
    sink_->write_data(vd_zero());   // :25 first data point is NULL
    if (!payload_) {                // :26-29 binary: one white voxel all leaves refer to
      VoxelData w = vd_zero();
      w.color = mk3(1.0f, 1.0f, 1.0f);
      sink_->write_data(w);
    }
  }

  // OctreeBuilder.cpp:131-146
  void add(uint64_t code) {
    if (code != cur_) fast_add_empty(code - cur_);
    Node n;
    n.data = 1;
    buf_[D_].push_back(n);
    refine(D_);
    cur_++;
  }
  // OctreeBuilder.cpp:149-168
  void add(const VoxelData& v) {
    if (v.morton != cur_) fast_add_empty(v.morton - cur_);
    Node n;
    n.data = sink_->write_data(v);
    n.cache = v;
    buf_[D_].push_back(n);
    refine(D_);
    cur_++;
  }

  // OctreeBuilder.cpp:34-55. The reference pads only while cur < max; cur == max (last
  // voxel at max-1) leaves the root buffer empty and reads it anyway (undefined). Here
  // the padding also runs in that case and hazard() records it.
  void finalize() {
    if (cur_ < max_) {
      fast_add_empty((max_ - cur_) + 1);
    } else if (cur_ == max_) {
      hazard_ = 1;
      fast_add_empty(1);
    }
    sink_->write_node(buf_[0][0]);
  }
  int hazard() const { return hazard_; }

 private:
  // OctreeBuilder.h:60-70: deepest level whose buffer is non-empty (-1 if none)
  int highest_non_empty() const {
    for (int k = D_; k >= 0; k--)
      if (!buf_[k].empty()) return k;
    return -1;
  }
  // OctreeBuilder.h:73-80
  int best_fill(uint64_t budget) const {
    int s = D_ - (int)find_power_of_8(budget);
    if (s == D_) return D_;
    return std::max(s, highest_non_empty());
  }
  // OctreeBuilder.h:83-91 (pow(8.0, .) is exact in double for the grids used here)
  void fast_add_empty(uint64_t budget) {
    while (budget > 0) {
      int b = best_fill(budget);
      add_empty(b);
      budget -= 1ull << (3 * (D_ - b));
    }
  }
  // OctreeBuilder.cpp:105-109
  void add_empty(int level) {
    buf_[level].push_back(Node());
    refine(level);
    cur_ += 1ull << (3 * (D_ - level));
  }
  // OctreeBuilder.cpp:112-128
  void refine(int start) {
    for (int d = start; d >= 0; d--) {
      if (buf_[d].size() == 8) {
        bool empty = true;
        for (int k = 0; k < 8; k++)
          if (!buf_[d][k].is_null()) empty = false;
        buf_[d - 1].push_back(empty ? Node() : group(buf_[d]));
        buf_[d].clear();
      } else {
        break;
      }
    }
  }
  // OctreeBuilder.cpp:58-102: write the non-null children consecutively in k order.
  Node group(const std::vector<Node>& b) {
    Node p;
    bool first = true;
    for (int k = 0; k < 8; k++) {
      if (!b[k].is_null()) {
        if (first) {
          p.children_base = sink_->write_node(b[k]);
          p.off[k] = 0;
          first = false;
        } else {
          p.off[k] = (int8_t)(sink_->write_node(b[k]) - p.children_base);
        }
      } else {
        p.off[k] = -1;
      }
    }
    if (levels_) {  // :82-99 average the 8 caches (null ones are zero) and store
      VoxelData d = vd_zero();
      float notnull = 0.0f;
      for (int i = 0; i < 8; i++) {
        if (!b[i].is_null()) notnull++;
        d.color = d.color + b[i].cache.color;
        d.normal = d.normal + b[i].cache.normal;
      }
      d.color = d.color / notnull;
      vec3 t = d.normal / notnull;
      d.normal = normalize(t);
      p.data = sink_->write_data(d);
      p.cache = d;
    }
    return p;
  }

  Sink* sink_;
  size_t G_;
  bool payload_, levels_;
  int D_ = 0;
  uint64_t cur_ = 0, max_ = 0;
  std::vector<std::vector<Node>> buf_;
  int hazard_ = 0;
};

// ------------------------------------------------------------------------------------
// Voxelizer: voxelizer.cpp:138-307 (Schwarz-Seidel conservative triangle/box overlap)
// with BarycentricCoords.h:4-32 for the payload colour.
// ------------------------------------------------------------------------------------
struct VoxOut {
  std::vector<uint64_t> codes;     // binary: emitted Morton codes
  std::vector<VoxelData> records;  // payload: emitted voxel data
  bool use_data = true;
  size_t nfilled = 0;
};

// BarycentricCoords.h:19-32 then :10-16 and :4-7
static inline vec3 barycentric_color(const Tri& t, vec3 n, float vx, float vy, float vz) {
  vec3 voxel = mk3(vx, vy, vz);
  float coeffD = -dot(t.v0, n);
  float k = (dot(voxel, n) + coeffD) / ((n.x * n.x + n.y * n.y) + n.z * n.z);
  vec3 point = voxel - k * n;
  mat3 M;
  M.c[0] = t.v0;
  M.c[1] = t.v1;
  M.c[2] = t.v2;
  vec3 b = mul(inverse(M), point);
  return (b.x * t.c0 + b.y * t.c1) + b.z * t.c2;
}

static void voxelize(const float* tris, size_t n, bool payload, uint64_t morton_start,
                     uint64_t morton_end, float unitlength, uint8_t* voxels,
                     float sparseness_limit, VoxOut& out) {
  const size_t stride = payload ? 21 : 9;
  std::memset(voxels, 0, (size_t)(morton_end - morton_start));  // :144
  out.codes.clear();
  out.records.clear();
  // :149-150 partition box in grid coordinates
  uint64_t pmn[3], pmx[3];
  morton_decode(morton_start, pmn[0], pmn[1], pmn[2]);
  morton_decode(morton_end - 1, pmx[0], pmx[1], pmx[2]);
  // :153-161 side-array capacity (line 159 overrides line 158 with sizeof(VoxelData))
  size_t data_max_items = 0;
  if (!payload && out.use_data) {
    uint64_t max_bytes = (uint64_t)((float)(morton_end - morton_start) * sparseness_limit);
    data_max_items = max_bytes / sizeof(VoxelData);
  }
  const float unit_div = 1.0f / unitlength;  // :164
  const vec3 delta_p = mk3(unitlength, unitlength, unitlength);

  for (size_t ti = 0; ti < n; ti++) {
    const Tri t = load_tri(tris + ti * stride, payload);
    if (!payload && out.use_data && out.codes.size() > data_max_items) out.use_data = false;

    // :189-204 bbox in world, truncated to grid, clamped to the partition
    float mn[3], mx[3];
    for (int a = 0; a < 3; a++) {
      mn[a] = fmin_std(t.v0[a], fmin_std(t.v1[a], t.v2[a]));
      mx[a] = fmax_std(t.v0[a], fmax_std(t.v1[a], t.v2[a]));
    }
    int gmn[3], gmx[3];
    for (int a = 0; a < 3; a++) {
      gmn[a] = clampval(trunc_i32_x86(mn[a] * unit_div), (int)pmn[a], (int)pmx[a]);
      gmx[a] = clampval(trunc_i32_x86(mx[a] * unit_div), (int)pmn[a], (int)pmx[a]);
    }

    // :207-254 per-triangle constants
    const vec3 e0 = t.v1 - t.v0;
    const vec3 e1 = t.v2 - t.v1;
    const vec3 e2 = t.v0 - t.v2;
    const vec3 nrm = normalize(cross(e0, e1));
    vec3 c = mk3(0.0f, 0.0f, 0.0f);
    if (nrm.x > 0) c.x = unitlength;
    if (nrm.y > 0) c.y = unitlength;
    if (nrm.z > 0) c.z = unitlength;
    const float d1 = dot(nrm, c - t.v0);
    const float d2 = dot(nrm, (delta_p - c) - t.v0);

    const vec3 E[3] = {e0, e1, e2};
    const vec3 V[3] = {t.v0, t.v1, t.v2};
    vec2 nxy[3], nyz[3], nzx[3];
    float dxy[3], dyz[3], dzx[3];
    for (int k = 0; k < 3; k++) {
      nxy[k] = vec2{-1.0f * E[k].y, E[k].x};
      nyz[k] = vec2{-1.0f * E[k].z, E[k].y};
      nzx[k] = vec2{-1.0f * E[k].x, E[k].z};
    }
    if (nrm.z < 0.0f)
      for (int k = 0; k < 3; k++) nxy[k] = -1.0f * nxy[k];
    if (nrm.x < 0.0f)
      for (int k = 0; k < 3; k++) nyz[k] = -1.0f * nyz[k];
    if (nrm.y < 0.0f)
      for (int k = 0; k < 3; k++) nzx[k] = -1.0f * nzx[k];
    for (int k = 0; k < 3; k++) {
      dxy[k] = ((-1.0f * dot(nxy[k], vec2{V[k].x, V[k].y})) + fmax_std(0.0f, unitlength * nxy[k].x)) +
               fmax_std(0.0f, unitlength * nxy[k].y);
      dyz[k] = ((-1.0f * dot(nyz[k], vec2{V[k].y, V[k].z})) + fmax_std(0.0f, unitlength * nyz[k].x)) +
               fmax_std(0.0f, unitlength * nyz[k].y);
      dzx[k] = ((-1.0f * dot(nzx[k], vec2{V[k].z, V[k].x})) + fmax_std(0.0f, unitlength * nzx[k].x)) +
               fmax_std(0.0f, unitlength * nzx[k].y);
    }

    // :257-303 test every voxel of the clamped bbox, x -> y -> z
    for (int x = gmn[0]; x <= gmx[0]; x++) {
      for (int y = gmn[1]; y <= gmx[1]; y++) {
        for (int z = gmn[2]; z <= gmx[2]; z++) {
          const uint64_t index = morton_encode((uint64_t)x, (uint64_t)y, (uint64_t)z);
          if (voxels[index - morton_start] == 1) continue;  // first triangle wins
          const vec3 p = mk3((float)x * unitlength, (float)y * unitlength, (float)z * unitlength);
          const float nDOTp = dot(nrm, p);
          if ((nDOTp + d1) * (nDOTp + d2) > 0.0f) continue;
          const vec2 pxy{p.x, p.y}, pyz{p.y, p.z}, pzx{p.z, p.x};
          bool pass = true;
          for (int k = 0; k < 3 && pass; k++)
            if ((dot(nxy[k], pxy) + dxy[k]) < 0.0f) pass = false;
          for (int k = 0; k < 3 && pass; k++)
            if ((dot(nyz[k], pyz) + dyz[k]) < 0.0f) pass = false;
          for (int k = 0; k < 3 && pass; k++)
            if ((dot(nzx[k], pzx) + dzx[k]) < 0.0f) pass = false;
          if (!pass) continue;
          voxels[index - morton_start] = 1;
          if (!payload) {
            if (out.use_data) out.codes.push_back(index);
          } else {
            vec3 col = barycentric_color(t, nrm, (float)x / unit_div, (float)y / unit_div,
                                         (float)z / unit_div);
            out.records.push_back(VoxelData{index, col, t.normal});
            out.codes.push_back(index);
          }
          out.nfilled++;
        }
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// Partitioner: partitioner.cpp:12-28 (estimate), :43-77 (boxes), :101-149 with
// BBoxBuffer.h:70-84 and intersection.h:9-18,49-53 (inclusive box test).
// ------------------------------------------------------------------------------------
static size_t estimate_partitions(size_t gridsize, size_t memory_limit, bool talk) {
  if (talk) std::cout << "Estimating best partition count ..." << std::endl;
  uint64_t required = (gridsize * gridsize * gridsize) / 1024 / 1024;
  if (talk)
    std::cout << "  to do this in-core I would need " << required << " Mb of system memory"
              << std::endl;
  if (required <= memory_limit) {
    if (talk) std::cout << "  memory limit of " << memory_limit << " Mb allows that" << std::endl;
    return 1;
  }
  size_t np = 1, req = required;
  while (req > memory_limit) {
    req = req / 8;
    np = np * 8;
  }
  if (talk)
    std::cout << "  going to do it in " << np << " partitions of " << req << " Mb each."
              << std::endl;
  return np;
}

struct Box {
  vec3 mn, mx;
};
static std::vector<Box> partition_boxes(float unitlength, size_t gridsize, size_t np) {
  std::vector<Box> boxes(np);
  const uint64_t morton_part = (gridsize * gridsize * gridsize) / np;
  for (size_t i = 0; i < np; i++) {
    uint64_t a[3], b[3];
    morton_decode(morton_part * i, a[0], a[1], a[2]);
    morton_decode(morton_part * (i + 1) - 1, b[0], b[1], b[2]);
    for (int k = 0; k < 3; k++) {
      boxes[i].mn[k] = (float)(uint32_t)a[k] * unitlength;
      boxes[i].mx[k] = (float)(uint32_t)(b[k] + 1) * unitlength;
    }
  }
  return boxes;
}
static inline bool box_overlap(const Box& a, const Box& b) {
  if (a.mx.x < b.mn.x || a.mx.y < b.mn.y || a.mx.z < b.mn.z || a.mn.x > b.mx.x ||
      a.mn.y > b.mx.y || a.mn.z > b.mx.z)
    return false;
  return true;
}
static inline Box tri_bbox(const float* p) {
  Box b;
  for (int a = 0; a < 3; a++) {
    b.mn[a] = fmin_std(p[a], fmin_std(p[3 + a], p[6 + a]));
    b.mx[a] = fmax_std(p[a], fmax_std(p[3 + a], p[6 + a]));
  }
  return b;
}
// per partition, the in-order list of triangle indices whose bbox touches its box
static std::vector<std::vector<uint64_t>> partition_lists(const float* tris, size_t n,
                                                          bool payload, float unitlength,
                                                          size_t gridsize, size_t np) {
  const size_t stride = payload ? 21 : 9;
  std::vector<Box> boxes = partition_boxes(unitlength, gridsize, np);
  std::vector<std::vector<uint64_t>> lists(np);
  for (size_t t = 0; t < n; t++) {
    Box b = tri_bbox(tris + t * stride);
    for (size_t j = 0; j < np; j++)
      if (box_overlap(b, boxes[j])) lists[j].push_back(t);
  }
  return lists;
}

// ------------------------------------------------------------------------------------
// Headers: tri_tools.h:75-127, trip_tools.h:56-125, octree_io.h:74-83 (iostream text;
// floats printed with the default 6 significant digits).
// ------------------------------------------------------------------------------------
struct TriInfo {
  std::string base;
  int version = 1;
  int geo_only = 0;
  size_t n_triangles = 0;
  float bbox[6] = {0, 0, 0, 0, 0, 0};
};
struct TripInfo {
  std::string base;
  int version = 1;
  int geo_only = 0;
  size_t gridsize = 0, n_triangles = 0, n_partitions = 0;
  float bbox[6] = {0, 0, 0, 0, 0, 0};
  std::vector<size_t> counts;
};

static int parse_tri_header(const std::string& fn, TriInfo& t) {
  std::ifstream f(fn.c_str());
  if (!f) {
    std::cout << "  Error: file " << fn << " does not exist." << std::endl;
    return 0;
  }
  t.base = fn.substr(0, fn.find_last_of("."));
  std::string line;
  f >> line;
  if (line != "#tri") {
    std::cout << "  Error: first line reads [" << line << "] instead of [#tri]" << std::endl;
    return 0;
  }
  f >> t.version;
  bool done = false;
  t.geo_only = 0;
  while (f.good() && !done) {
    f >> line;
    if (line == "END")
      done = true;
    else if (line == "ntriangles")
      f >> t.n_triangles;
    else if (line == "geo_only")
      f >> t.geo_only;
    else if (line == "bbox")
      f >> t.bbox[0] >> t.bbox[1] >> t.bbox[2] >> t.bbox[3] >> t.bbox[4] >> t.bbox[5];
    else {
      std::cout << "  unrecognized keyword [" << line << "], skipping" << std::endl;
      char c;
      do {
        c = (char)f.get();
      } while (f.good() && c != '\n');
    }
  }
  if (!done) {
    std::cout << "  error reading header" << std::endl;
    return 0;
  }
  return 1;
}

static void write_trip_header(const std::string& fn, const TripInfo& t, bool payload) {
  std::ofstream o(fn.c_str());
  o << "#trip " << t.version << std::endl;
  o << "gridsize " << t.gridsize << std::endl;
  o << "n_triangles " << t.n_triangles << std::endl;
  o << "bbox  " << t.bbox[0] << " " << t.bbox[1] << " " << t.bbox[2] << " " << t.bbox[3] << " "
    << t.bbox[4] << " " << t.bbox[5] << std::endl;
  o << "geo_only " << (payload ? 0 : 1) << std::endl;
  o << "n_partitions " << t.n_partitions << std::endl;
  for (size_t i = 0; i < t.n_partitions; i++) o << i << " " << t.counts[i] << std::endl;
  o << "END" << std::endl;
}

static int parse_trip_header(const std::string& fn, TripInfo& t) {
  std::ifstream f(fn.c_str());
  t.base = fn.substr(0, fn.find_last_of("."));
  std::string line;
  f >> line;
  if (line != "#trip") {
    std::cout << "  Error: first line reads [" << line << "] instead of [#trip]" << std::endl;
    return 0;
  }
  f >> t.version;
  bool done = false;
  t.geo_only = 0;
  while (f.good() && !done) {
    f >> line;
    if (line == "END")
      done = true;
    else if (line == "gridsize")
      f >> t.gridsize;
    else if (line == "n_triangles")
      f >> t.n_triangles;
    else if (line == "geo_only")
      f >> t.geo_only;
    else if (line == "bbox")
      f >> t.bbox[0] >> t.bbox[1] >> t.bbox[2] >> t.bbox[3] >> t.bbox[4] >> t.bbox[5];
    else if (line == "n_partitions") {
      f >> t.n_partitions;
      t.counts.assign(t.n_partitions, 0);
      for (size_t i = 0; i < t.n_partitions; i++) {
        int idx;
        size_t cnt;
        f >> idx >> cnt;
        if (idx >= 0 && (size_t)idx < t.n_partitions) t.counts[idx] = cnt;
      }
    } else {
      std::cout << "  unrecognized keyword [" << line << "], skipping" << std::endl;
      char c;
      do {
        c = (char)f.get();
      } while (f.good() && c != '\n');
    }
  }
  if (!done) {
    std::cout << "  error reading header" << std::endl;
    return 0;
  }
  return 1;
}

static void write_octree_header(const std::string& fn, size_t G, uint64_t nn, uint64_t nd) {
  std::ofstream o(fn.c_str());
  o << "#octreeheader 1" << std::endl;
  o << "gridlength " << G << std::endl;
  o << "n_nodes " << nn << std::endl;
  o << "n_data " << nd << std::endl;
  o << "END" << std::endl;
}

// The 6-significant-digit round trip the bbox takes through the .trip text header
// (trip_tools.h:112 print, main.cpp:306-307 re-parse).
static float text_roundtrip(float v) {
  std::ostringstream o;
  o << v;
  std::istringstream i(o.str());
  float r = 0.0f;
  i >> r;
  return r;
}

static bool file_exists(const std::string& n) {
  FILE* f = std::fopen(n.c_str(), "r");
  if (!f) return false;
  std::fclose(f);
  return true;
}
static std::string val_to_string(size_t v) {
  std::ostringstream o;
  o << v;
  return o.str();
}

// Colour modes applied after the sort: main.cpp:373-382
static void apply_color_mode(std::vector<VoxelData>& recs, int mode, size_t G) {
  for (auto& r : recs) {
    if (mode == 1) {
      r.color = mk3(1.0f, 1.0f, 1.0f);
    } else if (mode == 2) {
      r.color = morton_to_rgb(r.morton, G);
    } else if (mode == 3) {
      vec3 n = normalize(r.normal);
      r.color = mk3((n.x + 1.0f) / 2.0f, (n.y + 1.0f) / 2.0f, (n.z + 1.0f) / 2.0f);
    }
  }
}

// One partition's voxelize + sort + feed, as main.cpp:329-386 does.
static void run_partition(Builder& b, const float* tris, size_t n, bool payload,
                          uint64_t start, uint64_t end, float unitlength, uint8_t* voxels,
                          float sparseness, int color_mode, size_t G, size_t& nfilled) {
  VoxOut vo;
  vo.use_data = true;
  voxelize(tris, n, payload, start, end, unitlength, voxels, sparseness, vo);
  nfilled += vo.nfilled;
  if (!payload) {
    if (vo.use_data) {
      std::sort(vo.codes.begin(), vo.codes.end());
      for (uint64_t c : vo.codes) b.add(c);
    } else {  // side array overflowed: scan the grid (main.cpp:362-368)
      for (uint64_t j = 0; j < end - start; j++)
        if (voxels[j] != 0) b.add(start + j);
    }
  } else {
    std::sort(vo.records.begin(), vo.records.end(),
              [](const VoxelData& a, const VoxelData& c) { return a.morton < c.morton; });
    apply_color_mode(vo.records, color_mode, G);
    for (const VoxelData& r : vo.records) b.add(r);
  }
}

struct BuildResult {
  Sink sink;
  int hazard = 0;
};

}  // namespace orc

using namespace orc;

// ======================================================================================
// C interface
// ======================================================================================
extern "C" {

void oracle_morton_encode_batch(const uint32_t* x, const uint32_t* y, const uint32_t* z,
                                size_t n, uint64_t* out) {
  for (size_t i = 0; i < n; i++) out[i] = morton_encode(x[i], y[i], z[i]);
}

void oracle_morton_decode_batch(const uint64_t* m, size_t n, uint32_t* x, uint32_t* y,
                                uint32_t* z) {
  for (size_t i = 0; i < n; i++) {
    uint64_t a, b, c;
    morton_decode(m[i], a, b, c);
    x[i] = (uint32_t)a;
    y[i] = (uint32_t)b;
    z[i] = (uint32_t)c;
  }
}

size_t oracle_estimate_partitions(size_t gridsize, size_t memory_limit_mb) {
  return estimate_partitions(gridsize, memory_limit_mb, false);
}

void oracle_partition(const float* tris, size_t n, int payload, const float bbox[6],
                      size_t gridsize, size_t n_partitions, uint64_t* counts,
                      uint64_t* lists) {
  const float unitlength = (bbox[3] - bbox[0]) / (float)gridsize;  // partitioner.cpp:45
  auto L = partition_lists(tris, n, payload != 0, unitlength, gridsize, n_partitions);
  size_t o = 0;
  for (size_t j = 0; j < n_partitions; j++) {
    counts[j] = L[j].size();
    if (lists)
      for (uint64_t t : L[j]) lists[o++] = t;
  }
}

struct VoxHandle {
  VoxOut out;
};

void* oracle_voxelize(const float* tris, size_t n, int payload, uint64_t morton_start,
                      uint64_t morton_end, float unitlength, float sparseness_limit) {
  VoxHandle* h = new VoxHandle();
  std::vector<uint8_t> grid((size_t)(morton_end - morton_start));
  h->out.use_data = true;
  voxelize(tris, n, payload != 0, morton_start, morton_end, unitlength, grid.data(),
           sparseness_limit, h->out);
  if (!payload && !h->out.use_data) {  // overflowed: recover the set from the grid
    h->out.codes.clear();
    for (uint64_t j = 0; j < morton_end - morton_start; j++)
      if (grid[j]) h->out.codes.push_back(morton_start + j);
  }
  return h;
}
size_t oracle_vox_count(void* h) { return ((VoxHandle*)h)->out.codes.size(); }
int oracle_vox_used_data(void* h) { return ((VoxHandle*)h)->out.use_data ? 1 : 0; }
void oracle_vox_codes(void* h, uint64_t* out) {
  auto& c = ((VoxHandle*)h)->out.codes;
  if (!c.empty()) std::memcpy(out, c.data(), c.size() * 8);
}
void oracle_vox_records(void* h, void* out32) {
  auto& r = ((VoxHandle*)h)->out.records;
  uint8_t* o = (uint8_t*)out32;
  for (size_t i = 0; i < r.size(); i++) {
    float f[6] = {canon_nan(r[i].color.x),  canon_nan(r[i].color.y),  canon_nan(r[i].color.z),
                  canon_nan(r[i].normal.x), canon_nan(r[i].normal.y), canon_nan(r[i].normal.z)};
    std::memcpy(o + 32 * i, &r[i].morton, 8);
    std::memcpy(o + 32 * i + 8, f, 24);
  }
}
void oracle_vox_free(void* h) { delete (VoxHandle*)h; }

void* oracle_build(const uint64_t* codes, const void* records32, size_t n, size_t gridsize,
                   int payload, int levels) {
  BuildResult* r = new BuildResult();
  Builder b(&r->sink, gridsize, payload != 0, levels != 0);
  if (!payload) {
    for (size_t i = 0; i < n; i++) b.add(codes[i]);
  } else {
    const uint8_t* p = (const uint8_t*)records32;
    for (size_t i = 0; i < n; i++) {
      VoxelData v;
      std::memcpy(&v, p + 32 * i, 32);
      b.add(v);
    }
  }
  b.finalize();
  r->hazard = b.hazard();
  return r;
}
size_t oracle_build_nnodes(void* h) { return ((BuildResult*)h)->sink.node_pos; }
size_t oracle_build_ndata(void* h) { return ((BuildResult*)h)->sink.data_pos; }
int oracle_build_hazard(void* h) { return ((BuildResult*)h)->hazard; }
void oracle_build_copy(void* h, void* nodes_out, void* data_out) {
  BuildResult* r = (BuildResult*)h;
  if (nodes_out && !r->sink.nodes.empty())
    std::memcpy(nodes_out, r->sink.nodes.data(), r->sink.nodes.size());
  if (data_out && !r->sink.data.empty())
    std::memcpy(data_out, r->sink.data.data(), r->sink.data.size());
}
void oracle_build_free(void* h) { delete (BuildResult*)h; }

void* oracle_pipeline(const float* tris, size_t n, int payload, const float bbox[6],
                      size_t gridsize, size_t memory_limit_mb, float sparseness_limit,
                      int color_mode, int levels, size_t* n_partitions_out,
                      size_t* nfilled_out) {
  const bool pl = payload != 0;
  const size_t stride = pl ? 21 : 9;
  const size_t np = estimate_partitions(gridsize, memory_limit_mb, false);
  // partitioner.cpp:45 uses the .tri bbox; main.cpp:311 the bbox re-read from .trip text
  const float u_part = (bbox[3] - bbox[0]) / (float)gridsize;
  const float u_vox = (text_roundtrip(bbox[3]) - text_roundtrip(bbox[0])) / (float)gridsize;
  std::vector<std::vector<uint64_t>> lists;
  if (np > 1) lists = partition_lists(tris, n, pl, u_part, gridsize, np);
  const uint64_t morton_part = (gridsize * gridsize * gridsize) / np;
  // main.cpp:314 allocates the grid uninitialised; voxelize clears it per partition
  std::unique_ptr<uint8_t[]> voxels(new uint8_t[(size_t)morton_part]);
  BuildResult* r = new BuildResult();
  Builder b(&r->sink, gridsize, pl, levels != 0);
  size_t nfilled = 0;
  std::vector<float> sub;
  for (size_t i = 0; i < np; i++) {
    const float* src = tris;
    size_t cnt = n;
    if (np > 1) {
      cnt = lists[i].size();
      sub.resize(cnt * stride);
      for (size_t k = 0; k < cnt; k++)
        std::memcpy(&sub[k * stride], tris + lists[i][k] * stride, stride * 4);
      src = sub.data();
    }
    if (cnt == 0) continue;  // main.cpp:330
    run_partition(b, src, cnt, pl, i * morton_part, (i + 1) * morton_part, u_vox,
                  voxels.get(), sparseness_limit, color_mode, gridsize, nfilled);
  }
  b.finalize();
  r->hazard = b.hazard();
  if (n_partitions_out) *n_partitions_out = np;
  if (nfilled_out) *nfilled_out = nfilled;
  return r;
}

// main.cpp:100-399 on files, with the same flags, messages and exit statuses.
int oracle_cli_main(int argc, char** argv, int payload) {
  const bool pl = payload != 0;
  using std::cout;
  using std::endl;
  std::string filename;
  size_t gridsize = 1024, memlimit = 2048;
  float sparseness = 0.10f;
  int color = 0;
  bool levels = false, verbose = false;
  auto help = [] {
    cout << "Example: svo_builder -f /home/jeroen/bunny.tri" << endl << endl;
    cout << "All available program options:" << endl << endl;
    cout << "-f <filename.tri>     Path to a .tri input file." << endl;
    cout << "-s <gridsize>         Voxel gridsize, should be a power of 2. Default 512." << endl;
    cout << "-l <memory_limit>     Memory limit for process, in Mb. Default 1024." << endl;
    cout << "-levels               Generate intermediary voxel levels by averaging voxel data" << endl;
    cout << "-c <option>           Coloring of voxels (Options: model (default), fixed, linear, normal)" << endl;
    cout << "-d <percentage>       Percentage of memory limit to be used additionaly for sparseness optimization" << endl;
    cout << "-v                    Be very verbose." << endl;
    cout << "-h                    Print help and exit." << endl;
  };
  auto invalid = [&] {
    cout << "Not enough or invalid arguments, please try again." << endl;
    cout << "At the bare minimum, I need a path to a .TRI file" << endl << endl;
    help();
  };
  cout << "Reading program parameters ..." << endl;
  if (argc < 3) {
    invalid();
    return 0;
  }
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    bool has_next = i + 1 < argc;
    if (a == "-f" && has_next) {
      filename = argv[++i];
      if (filename.find(".tri") == std::string::npos) {
        cout << "Data filename does not end in .tri - I only support that file format" << endl;
        invalid();
        return 0;
      }
    } else if (a == "-s" && has_next) {
      gridsize = (size_t)atoi(argv[++i]);
      if (!is_pow2((unsigned)gridsize)) {
        cout << "Requested gridsize is not a power of 2" << endl;
        invalid();
        return 0;
      }
    } else if (a == "-l" && has_next) {
      memlimit = (size_t)atoi(argv[++i]);
      if (memlimit <= 1) {
        cout << "Requested memory limit is nonsensical. Use a value >= 1" << endl;
        invalid();
        return 0;
      }
    } else if (a == "-d" && has_next) {
      int percent = atoi(argv[++i]);
      sparseness = percent / 100.0f;
      if (sparseness < 0) {
        cout << "Requested data memory limit is nonsensical. Use a value > 0" << endl;
        invalid();
        return 0;
      }
    } else if (a == "-v") {
      verbose = true;
    } else if (a == "-levels") {
      levels = true;
    } else if (a == "-c" && has_next) {
      std::string c = argv[++i];
      if (!pl) {
        cout << "You asked to generate colors, but we're only doing binary voxelisation." << endl;
      } else if (c == "model") {
        color = 0;
      } else if (c == "linear") {
        color = 2;
      } else if (c == "normal") {
        color = 3;
      } else if (c == "fixed") {
        color = 1;
      } else {
        cout << "Unrecognized color switch: " << c << ", so reverting to colors from model." << endl;
      }
    } else if (a == "-h") {
      help();
      return 0;
    } else {
      invalid();
      return 0;
    }
  }
  (void)verbose;

  TriInfo ti;
  cout << "Parsing tri header " << filename << " ..." << endl;
  if (parse_tri_header(filename, ti) != 1) return 0;
  if (!(file_exists(ti.base + ".tri") && file_exists(ti.base + ".tridata"))) {
    cout << "Not all required .tri or .tridata files exist. Please regenerate using tri_convert." << endl;
    return 0;
  }
  if (!pl && !ti.geo_only) {
    cout << "You're using a .tri file which contains more than just geometry with a geometry-only SVO Builder! Regenerate that .tri file using tri_convert_binary." << endl;
    return 0;
  }
  if (pl && ti.geo_only) {
    cout << "You're using a .tri file which contains only geometry with the regular SVO Builder! Regenerate that .tri file using tri_convert." << endl;
    return 0;
  }
  const size_t stride = pl ? 21 : 9;
  // whole .tridata into memory; short files are an error
  std::vector<float> tris(ti.n_triangles * stride);
  {
    FILE* f = std::fopen((ti.base + ".tridata").c_str(), "rb");
    size_t got = f ? std::fread(tris.data(), stride * 4, ti.n_triangles, f) : 0;
    if (f) std::fclose(f);
    if (got != ti.n_triangles) {
      cout << "Error: " << ti.base << ".tridata holds " << got << " triangles, header says "
           << ti.n_triangles << endl;
      return 1;
    }
  }
  const size_t np = estimate_partitions(gridsize, memlimit, true);
  cout << "Partitioning data into " << np << " partitions ... ";
  cout.flush();
  TripInfo tp;
  tp.version = ti.version;
  tp.gridsize = gridsize;
  tp.n_triangles = ti.n_triangles;
  tp.n_partitions = np;
  std::memcpy(tp.bbox, ti.bbox, sizeof(tp.bbox));
  tp.base = ti.base + val_to_string(gridsize) + "_" + val_to_string(np);
  const float u_part = (ti.bbox[3] - ti.bbox[0]) / (float)gridsize;
  std::vector<std::vector<uint64_t>> lists;
  if (np > 1) {
    lists = partition_lists(tris.data(), ti.n_triangles, pl, u_part, gridsize, np);
    for (size_t j = 0; j < np; j++) tp.counts.push_back(lists[j].size());
  } else {
    tp.counts.push_back(ti.n_triangles);
  }
  // .tripdata files, as the reference leaves them until the end (removed below)
  for (size_t j = 0; j < np; j++) {
    if (tp.counts[j] == 0) continue;
    FILE* f = std::fopen((tp.base + "_" + val_to_string(j) + ".tripdata").c_str(), "wb");
    if (np == 1) {
      std::fwrite(tris.data(), stride * 4, ti.n_triangles, f);
    } else {
      for (uint64_t t : lists[j]) std::fwrite(&tris[t * stride], stride * 4, 1, f);
    }
    std::fclose(f);
  }
  write_trip_header(tp.base + ".trip", tp, pl);
  cout << "done." << endl;

  TripInfo tr;
  if (parse_trip_header(tp.base + ".trip", tr) != 1) return 0;
  const float unitlength = (tr.bbox[3] - tr.bbox[0]) / (float)tr.gridsize;
  const uint64_t morton_part = (tr.gridsize * tr.gridsize * tr.gridsize) / tr.n_partitions;
  std::unique_ptr<uint8_t[]> voxels(new uint8_t[(size_t)morton_part]);  // main.cpp:314

  BuildResult r;
  r.sink.fn = std::fopen((tr.base + ".octreenodes").c_str(), "wb");
  r.sink.fd = std::fopen((tr.base + ".octreedata").c_str(), "wb");
  Builder b(&r.sink, tr.gridsize, pl, levels);
  size_t nfilled = 0;
  std::vector<float> part;
  for (size_t i = 0; i < tr.n_partitions; i++) {
    if (tr.counts[i] == 0) continue;
    cout << "Voxelizing partition " << i << " ..." << endl;
    part.resize(tr.counts[i] * stride);
    FILE* f = std::fopen((tr.base + "_" + val_to_string(i) + ".tripdata").c_str(), "rb");
    size_t got = f ? std::fread(part.data(), stride * 4, tr.counts[i], f) : 0;
    if (f) std::fclose(f);
    if (got != tr.counts[i]) {
      cout << "Error: short partition file" << endl;
      return 1;
    }
    cout << "Building SVO for partition " << i << " ..." << endl;
    run_partition(b, part.data(), tr.counts[i], pl, i * morton_part, (i + 1) * morton_part,
                  unitlength, voxels.get(), sparseness, color, tr.gridsize, nfilled);
  }
  b.finalize();
  if (b.hazard())
    std::cerr << "warning: the last filled voxel is the second-to-last Morton code; the "
                 "reference leaves the root undefined here, the completed tree is written"
              << endl;
  std::fclose(r.sink.fn);
  std::fclose(r.sink.fd);
  write_octree_header(tr.base + ".octree", tr.gridsize, r.sink.node_pos, r.sink.data_pos);
  cout << "done" << endl;
  cout << "Total amount of voxels: " << nfilled << endl;
  // partitioner.cpp:31-40
  std::remove((tr.base + ".trip").c_str());
  for (size_t i = 0; i < tr.n_partitions; i++)
    std::remove((tr.base + "_" + val_to_string(i) + ".tripdata").c_str());
  return 0;
}

}  // extern "C"

#ifdef ORACLE_MAIN
int main(int argc, char** argv) { return oracle_cli_main(argc, argv, ORACLE_PAYLOAD); }
#endif
