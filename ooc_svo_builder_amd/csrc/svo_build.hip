// Bottom-up SVO build from the sorted unique Morton codes, replacing the streaming
// OctreeBuilder (OctreeBuilder.cpp:34-168, OctreeBuilder.h:50-91) with level passes.
//
// Node order. The streaming builder writes, each time a node completes, its non-null
// children consecutively in child order (groupNodes, OctreeBuilder.cpp:58-79), and the
// root last (:42). Nodes complete in post-order, so the children written for the whole
// subtree of a node P occupy one contiguous range of size(P) - 1 records (size = nodes in
// the subtree, P included), laid out as: the ranges of P's children subtrees in child
// order, then P's own children. Hence, with start(P) the first record of P's range:
//   start(root)        = 0
//   start(child r)     = start(P) + sum_{r' < r} (size(child r') - 1)
//   children_base(P)   = start(P) + size(P) - 1 - popcount(mask(P))
//   position(child r)  = children_base(P) + r,   root at n_nodes - 1 = size(root) - 1.
// Bottom-up passes build each level (parent = code >> 3, child mask, subtree size) with
// one scan per level; top-down passes give every node its start and position and write
// its 24-byte record (Node.h:16-18, octree_io.h:62-66), one thread per node, level by
// level from the root.
#include <algorithm>
#include <cstring>
#include <vector>

#include "svo_device.h"
#include "svo_internal.h"
#include "svo_scan.h"

namespace svo {

typedef unsigned long long ull;

// child offsets of a node with child mask m: rank of each present child, -1 otherwise
__device__ __forceinline__ uint64_t offsets_of(uint32_t m) {
  uint64_t o = 0;
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint64_t byte = (m >> k & 1u) ? (uint64_t)(r++) : 0xFFull;
    o |= byte << (8 * k);
  }
  return o;
}

// A node's data index and children_base as written: moved by the job's placement in a
// larger node stream (zero but for a streamed partition), data indices only past the
// leading records (the null record, and the white record of geometry-only leaves).
__device__ __forceinline__ uint64_t placed_data(const JobDev& J, uint64_t data) {
  return data >= (J.payload ? 1ull : 2ull) ? data + J.data_add : data;
}

__device__ __forceinline__ void put_node(const JobDev& J, uint64_t* nodes, uint64_t pos,
                                         uint64_t data, uint64_t cb, uint64_t offs) {
  uint64_t* r = nodes + 3 * pos;
  r[0] = placed_data(J, data);
  r[1] = offs != ~0ull ? cb + J.cb_add : cb;
  r[2] = offs;
}

// Jobs. The voxel stream may hold several jobs (meshes), job-major, each tagged in the
// code bits selected by tag_mask (0 for a single job). Levels are indexed by height above
// the leaves, h = 0 (leaves) .. H: a node of job j at height h has depth D_j - h, so the
// root of job j is its only node at height D_j, and one pass per height serves all jobs.
// parent(c) keeps the tag and drops the lowest Morton digit.
struct Tags {
  uint64_t tag_mask;  // bits of the job index (shifted by kJobShift)
  __device__ __forceinline__ uint32_t job(uint64_t c) const {
    return (uint32_t)((c & tag_mask) >> kJobShift);
  }
  __device__ __forceinline__ uint64_t parent(uint64_t c) const {
    return (c & tag_mask) | ((c & ~tag_mask) >> 3);
  }
};

// Level arrays live in arenas (code, size, mask, first child, start, parent index),
// height h >= 1 at element offset off[h]; the per-height counts n[h] and offsets are kept
// on the device so that the passes run back to back without host round trips. Height 0
// (the leaves) is the voxel code stream itself, with implicit size 1.
struct LevelMeta {
  unsigned long long* n;    // n[h], h = 0..H
  unsigned long long* off;  // off[h], h = 1..H+1
  unsigned long long* flag; // [0] arena overflow, [1] node range error, [2] total nodes
};

struct Arena {
  uint64_t* code;
  uint64_t* size;
  uint8_t* mask;
  uint32_t* first;
  uint64_t* start;
  uint32_t* pidx;  // index of the node's parent within the parent level
  uint64_t cap;
  // -levels only (nullptr otherwise): internal nodes in the subtree (the node included),
  // the node's averaged colour and normal, and the first data-record rank of its subtree
  uint64_t* isize;
  float* cache;  // 6 floats per node
  uint64_t* pstart;
  // sum of (size - 1) over the node's preceding siblings: its start is its parent's start
  // plus this (written by the bottom-up pass that creates the parent)
  uint64_t* sib;
};

// -levels data record of an internal node (OctreeBuilder.cpp:82-99): the caches of its
// 8 children summed in child order (a null child adds zeros), divided by the number of
// non-null children, the normal then normalized (glm::normalize = v * (1 / sqrt(dot))).
// Children j in [i, je) of the child level, in child order; leaf caches are the voxels'
// own records (payload) or zero (geometry-only builds: leaves carry no data).
__device__ __forceinline__ void average_children(const Arena& A, uint64_t coff, bool leaf,
                                                 const float* __restrict__ leaf_rec,
                                                 int payload, const uint64_t* code, uint64_t i,
                                                 uint64_t je, uint64_t& isz, float* out) {
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
  float notnull = 0.0f;
  isz = 1;
  uint64_t j = i;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, b0 = 0.0f, b1 = 0.0f, b2 = 0.0f;
    if (j < je && (int)(code[j] & 7) == k) {
      notnull += 1.0f;
      const float* r = nullptr;
      if (!leaf) {
        isz += A.isize[coff + j];
        r = A.cache + (coff + j) * 6;
      } else if (payload) {
        r = leaf_rec + j * 8 + 2;
      }
      if (r) {
        a0 = r[0];
        a1 = r[1];
        a2 = r[2];
        b0 = r[3];
        b1 = r[4];
        b2 = r[5];
      }
      j++;
    }
    c0 = c0 + a0;
    c1 = c1 + a1;
    c2 = c2 + a2;
    n0 = n0 + b0;
    n1 = n1 + b1;
    n2 = n2 + b2;
  }
  out[0] = c0 / notnull;
  out[1] = c1 / notnull;
  out[2] = c2 / notnull;
  const float t0 = n0 / notnull, t1 = n1 / notnull, t2 = n2 / notnull;
  const float inv = 1.0f / sqrtf((t0 * t0 + t1 * t1) + t2 * t2);
  out[3] = t0 * inv;
  out[4] = t1 * inv;
  out[5] = t2 * inv;
}

// Parent height h+1 from child height h: head = first child of a parent group; a job's
// root (height D_j) has no parent.
struct HeadLoad {
  Arena A;
  LevelMeta M;
  Tags T;
  const JobDev* jobs;
  const uint64_t* leaves;
  int h;
  int pad = 0;  // no padding bytes: launch logs compare arguments byte by byte
  __device__ uint32_t operator()(uint64_t i) const {
    // an overflowed child level is not read (its tail lies beyond the arena)
    if (h > 0 && M.off[h] + M.n[h] > A.cap) return 0u;
    const uint64_t* code = (h == 0) ? leaves : A.code + M.off[h];
    const uint64_t c = code[i];
    const uint64_t cp = i ? code[i - 1] : 0ull;
    // one job (no tags): its depth does not wait for the code
    const int depth = T.tag_mask ? jobs[T.job(c)].depth : jobs[0].depth;
    if (depth <= h) return 0u;
    return (i == 0 || T.parent(c) != T.parent(cp)) ? 1u : 0u;
  }
};
struct ParentStore {
  Arena A;
  LevelMeta M;
  Tags T;
  const JobDev* jobs;
  const uint64_t* leaves;
  uint32_t* leaf_parent;  // height 0 only: parent index of every leaf
  unsigned long long* root_size;  // per job: nodes in its tree
  unsigned long long* root_isize; // per job: internal nodes in its tree (-levels)
  const float* leaf_rec;          // -levels, payload: voxel i's record at leaf_rec + 8 i
  int h;
  int pad = 0;  // no padding bytes: launch logs compare arguments byte by byte
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t head, bool) const {
    const bool leaf = (h == 0);
    if (leaf)
      leaf_parent[i] = ex + head - 1;
    else if (M.off[h] + i < A.cap)
      A.pidx[M.off[h] + i] = ex + head - 1;
    if (!leaf && M.off[h] + M.n[h] > A.cap) {
      M.flag[0] = 1;
      return;
    }
    const uint64_t* code = leaf ? leaves : A.code + M.off[h];
    const uint64_t* csize = leaf ? nullptr : A.size + M.off[h];
    const uint64_t nch = M.n[h];
    const uint64_t base = M.off[h + 1];
    if (head) {
      if (base + ex >= A.cap) {
        M.flag[0] = 1;
      } else {
        // the (at most 8) children follow the head: all their loads issued at once
        uint64_t cj[8], sj[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          cj[k] = (i + k < nch) ? code[i + k] : ~0ull;  // ~0: no node has this parent
          sj[k] = (csize && i + k < nch) ? csize[i + k] : 1ull;
        }
        const uint64_t pc = T.parent(cj[0]);
        uint32_t m = 0;
        uint64_t sz = 1, je = i;
        bool more = true;
#pragma unroll
        for (int k = 0; k < 8; k++) {
          more = more && T.parent(cj[k]) == pc;
          if (more) {
            m |= 1u << (cj[k] & 7);
            if (csize) A.sib[M.off[h] + i + k] = sz - 1 - (uint64_t)k;
            sz += sj[k];
            je = i + k + 1;
          }
        }
        A.code[base + ex] = pc;
        A.size[base + ex] = sz;
        A.mask[base + ex] = (uint8_t)m;
        A.first[base + ex] = (uint32_t)i;
        const uint32_t jb = T.job(pc);
        uint64_t isz = 1;
        if (A.isize) {
          float avg[6];
          average_children(A, leaf ? 0 : M.off[h], leaf, leaf_rec, jobs[jb].payload, code, i,
                           je, isz, avg);
          A.isize[base + ex] = isz;
#pragma unroll
          for (int q = 0; q < 6; q++) A.cache[(base + ex) * 6 + q] = avg[q];
        }
        if (jobs[jb].depth == h + 1) {
          root_size[jb] = sz;
          root_isize[jb] = isz;
        }
      }
    }
    if (i == nch - 1) {
      const uint64_t np = (uint64_t)ex + head;
      M.n[h + 1] = np;
      M.off[h + 2] = base + np;
    }
  }
};

// Level bookkeeping before a build: all zero, n[0] = V and, when the four lowest levels
// come from the brick stream, n[4] = NB and off[5] = NB.
__global__ __launch_bounds__(256) void k_build_init(unsigned long long* meta, uint64_t words,
                                                    uint64_t V, uint64_t NB, int fast) {
  for (uint64_t i = threadIdx.x; i < words; i += blockDim.x) {
    unsigned long long v = 0;
    if (i == 0) v = V;
    if (fast && (i == 4 || i == 64 + 5)) v = NB;
    meta[i] = v;
  }
}

// Per job, after the bottom-up passes: node count (a single null root for an empty
// grid, finalizeTree pads it; the single leaf of a 1^3 grid), node base (exclusive scan
// over the jobs), data count (-levels adds one record per internal node), the records
// the builder writes itself (OctreeBuilder.cpp:25-29: the null record, plus the white
// record of geometry-only builds), and the hazard flag. One block.
// BLOCK threads of one block; wsum holds BLOCK / 64 words.
template <int BLOCK>
__device__ __forceinline__ void job_bases(JobDev* jobs, uint32_t n_jobs,
                                          const unsigned long long* root_size,
                                          const unsigned long long* root_isize,
                                          const unsigned long long* brick_hazard,
                                          const uint64_t* __restrict__ codes, Tags T,
                                          LevelMeta M, uint64_t node_cap,
                                          uint64_t* __restrict__ nodes,
                                          const float* __restrict__ in_rec,
                                          float* __restrict__ records, uint64_t* wsum,
                                          uint64_t& carry) {
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t j0 = 0; j0 < n_jobs; j0 += BLOCK) {
    const uint32_t j = j0 + threadIdx.x;
    uint64_t nn = 0;
    if (j < n_jobs) {
      const JobDev& J = jobs[j];
      nn = (J.nvox == 0 || J.depth == 0) ? 1ull : root_size[j];
    }
    uint64_t tot;
    const uint64_t inc = block_incl_scan<uint64_t, BLOCK>(nn, wsum, tot);
    const uint64_t nb = carry + inc - nn;
    if (j < n_jobs) {
      JobDev& J = jobs[j];
      J.n_nodes = nn;
      J.nbase = nb;
      const bool trivial = (J.nvox == 0 || J.depth == 0);
      J.n_data = (J.payload ? 1 + J.nvox : 2) + ((J.levels && !trivial) ? root_isize[j] : 0);
      const uint64_t G3 = J.gridsize * J.gridsize * J.gridsize;
      uint64_t hz = 0;
      if (J.nvox == 0) {
        hz = (J.gridsize == 1) ? 1 : 0;
      } else if (brick_hazard) {  // found in the brick stream (codes not written)
        hz = brick_hazard[j];
      } else {
        const uint64_t last = codes[J.voff + J.nvox - 1] & ~T.tag_mask;
        // OctreeBuilder.cpp:36-42 leaves the root undefined when the last voxel is the
        // second-to-last code; the completed tree is built and the hazard reported
        hz = (G3 >= 2 && last == G3 - 2) ? 1 : 0;
      }
      J.hazard = hz;
      if (trivial) {
        if (nb < node_cap) {
          put_node(J, nodes, nb, J.nvox ? 1ull : 0ull, 0ull, ~0ull);
        } else {
          M.flag[1] = 1;
        }
      }
      float* d = records + J.dbase * 8;
      if (J.levels && J.payload && J.depth == 0 && J.nvox && in_rec != records) {
        // 1^3 grid: the single voxel's record keeps its place after the null record
        for (int k = 0; k < 8; k++) d[8 + k] = in_rec[(J.dbase + 1) * 8 + k];
      }
      for (int k = 0; k < 8; k++) d[k] = 0.0f;
      if (!J.payload) {
        // data[1] = {0, colour (1,1,1), normal (0,0,0)}
        for (int k = 0; k < 8; k++) d[8 + k] = 0.0f;
        d[8 + 2] = d[8 + 3] = d[8 + 4] = 1.0f;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) M.flag[2] = carry;
}

// The bottom-up levels h_lo .. H-1 whose element counts are small (bounded on the host),
// in one block, level after level (a level's scan in chunks of the block with a running
// carry, a barrier before the next level reads it), then the per-job bases.
constexpr int kTailBlock = 1024;
constexpr uint64_t kTailLevelMax = 8192;  // largest level bound the tail block takes
__global__ __launch_bounds__(kTailBlock) void k_levels_tail(
    Arena A, LevelMeta M, Tags T, JobDev* jobs, uint32_t n_jobs, const uint64_t* leaves,
    uint32_t* leaf_parent, unsigned long long* root_size, unsigned long long* root_isize,
    const float* leaf_rec, int h_lo, int H, const unsigned long long* brick_hazard,
    uint64_t node_cap, uint64_t* __restrict__ nodes, const float* __restrict__ in_rec,
    float* __restrict__ records) {
  __shared__ uint32_t wsum32[kTailBlock / kWave];
  __shared__ uint64_t wsum[kTailBlock / kWave];
  __shared__ uint64_t carry;
  for (int h = h_lo; h < H; h++) {
    HeadLoad ld{A, M, T, jobs, leaves, h};
    ParentStore st{A, M, T, jobs, leaves, leaf_parent, root_size, root_isize, leaf_rec, h};
    __syncthreads();  // the previous level's nodes and counts are written
    const uint64_t n = M.n[h];
    uint32_t run = 0;
    for (uint64_t b = 0; b < n; b += kTailBlock) {
      const uint64_t i = b + threadIdx.x;
      const uint32_t v = i < n ? ld(i) : 0u;
      uint32_t tot;
      const uint32_t inc = block_incl_scan<uint32_t, kTailBlock>(v, wsum32, tot);
      if (i < n) st(i, run + inc - v, v, i == n - 1);
      run += tot;
    }
  }
  __syncthreads();
  job_bases<kTailBlock>(jobs, n_jobs, root_size, root_isize, brick_hazard, leaves, T, M,
                        node_cap, nodes, in_rec, records, wsum, carry);
}

// -levels: the data records of a subtree are written in post-order (a node completes
// after its children, OctreeBuilder.cpp:58-99, 149-168), so node P's record is at
// base + pstart(P) + S(P) - 1, with S = nodes that write a record (every node for
// payload builds, internal nodes only for geometry-only ones) and
// pstart(child r) = pstart(P) + sum_{r' < r} S(child r').
__device__ __forceinline__ uint64_t levels_weight(const Arena& A, uint64_t k, int payload) {
  return payload ? A.size[k] : A.isize[k];
}

__device__ __forceinline__ void put_data(float* rec, uint64_t idx, const float* avg) {
  float* r = rec + idx * 8;
  *(uint64_t*)r = 0ull;  // internal nodes' records carry morton 0
#pragma unroll
  for (int q = 0; q < 6; q++) r[2 + q] = canon_nan(avg[q]);
}

// One thread per node at height h >= 1: a job's root starts at 0 and sits last; any
// other node's start is its parent's start plus the (size - 1) of its preceding
// siblings, and its position is the parent's children_base plus its rank. It writes its
// own record (children_base relative to its job's node array) and start; with -levels
// also its data record. Consecutive nodes land in consecutive records, so the stores
// coalesce.
__global__ __launch_bounds__(256) void k_emit_level(Arena A, LevelMeta M, Tags T,
                                                    const JobDev* __restrict__ jobs, int h,
                                                    uint64_t node_cap,
                                                    uint64_t* __restrict__ nodes,
                                                    float* __restrict__ out_rec) {
  if (M.flag[0]) return;  // arena overflow: the build is redone with a larger arena
  const uint64_t nc = M.n[h];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const uint64_t co = M.off[h];
  const uint64_t code = A.code[co + i];
  const JobDev& J = jobs[T.job(code)];
  const bool lv = A.isize && J.levels;
  const uint64_t sz = A.size[co + i];
  const uint32_t cm = A.mask[co + i];
  uint64_t st, pos, pst = 0;
  if (J.depth == h) {
    st = 0;
    pos = sz - 1;
  } else {
    const uint64_t po = M.off[h + 1] + A.pidx[co + i];
    const uint32_t f = A.first[po];
    const uint32_t pcnt = (uint32_t)__popc((uint32_t)A.mask[po]);
    const uint64_t ps = A.start[po];
    const uint64_t cb = ps + A.size[po] - 1 - pcnt;
    st = ps;
    for (uint64_t j = f; j < i; j++) st += A.size[co + j] - 1;
    if (lv) {
      pst = A.pstart[po];
      for (uint64_t j = f; j < i; j++) pst += levels_weight(A, co + j, J.payload);
    }
    pos = cb + (i - f);
  }
  const uint64_t gpos = J.nbase + pos;
  if (pos >= J.n_nodes || gpos >= node_cap) {
    M.flag[1] = 1;
    return;
  }
  A.start[co + i] = st;
  uint64_t data = 0;
  if (lv) {
    A.pstart[co + i] = pst;
    const uint64_t idx = (J.payload ? 1 : 2) + pst + levels_weight(A, co + i, J.payload) - 1;
    if (idx >= J.n_data) {
      M.flag[1] = 1;
      return;
    }
    data = idx;
    put_data(out_rec, J.dbase + idx, A.cache + (co + i) * 6);
  }
  put_node(J, nodes, gpos, data, st + sz - 1 - (uint32_t)__popc(cm), offsets_of(cm));
}

// Every node of heights hs..H in one launch (geometry records, no -levels data): one
// thread per node, flattened over the contiguous levels of the arena. A node's start is
// the sum of the sibling offsets along its path to the root (start(P) plus its own sib),
// which needs only the bottom-up results, so all levels are written at once.
__global__ __launch_bounds__(256) void k_emit_upper(Arena A, LevelMeta M, Tags T,
                                                    const JobDev* __restrict__ jobs, int hs,
                                                    int H, uint64_t node_cap,
                                                    uint64_t* __restrict__ nodes) {
  if (M.flag[0]) return;
  const uint64_t g = M.off[hs] + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int h = hs;
  while (h <= H && g >= M.off[h] + M.n[h]) h++;
  if (h > H) return;
  const uint64_t i = g - M.off[h];
  const uint64_t code = A.code[g];
  const JobDev& J = jobs[T.job(code)];
  const uint64_t sz = A.size[g];
  const uint32_t cm = A.mask[g];
  uint64_t st, pos;
  if (J.depth == h) {  // the job's root: start 0, last record
    st = 0;
    pos = sz - 1;
  } else {
    const uint64_t po = M.off[h + 1] + A.pidx[g];
    uint64_t sp = 0;  // start of the parent
    uint64_t cur = po;
    for (int l = h + 1; l < J.depth; l++) {
      sp += A.sib[cur];
      cur = M.off[l + 1] + A.pidx[cur];
    }
    st = sp + A.sib[g];
    pos = sp + A.size[po] - 1 - (uint32_t)__popc((uint32_t)A.mask[po]) + (i - A.first[po]);
  }
  const uint64_t gpos = J.nbase + pos;
  if (pos >= J.n_nodes || gpos >= node_cap) {
    M.flag[1] = 1;
    return;
  }
  A.start[g] = st;
  put_node(J, nodes, gpos, 0ull, st + sz - 1 - (uint32_t)__popc(cm), offsets_of(cm));
}

// One thread per leaf: position = children_base(parent) + rank among its siblings.
// OctreeBuilder.cpp:137-139 (binary: data 1) / :157-159 (payload: its data index; with
// -levels its record moves to its post-order place).
__global__ __launch_bounds__(256) void k_emit_leaves(Arena A, LevelMeta M, Tags T,
                                                     const JobDev* __restrict__ jobs,
                                                     uint64_t node_cap,
                                                     const uint64_t* __restrict__ codes,
                                                     const uint32_t* __restrict__ leaf_parent,
                                                     uint64_t* __restrict__ nodes,
                                                     const float* __restrict__ in_rec,
                                                     float* __restrict__ out_rec) {
  if (M.flag[0]) return;
  const uint64_t nl = M.n[0];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl) return;
  const JobDev& J = jobs[T.job(codes[i])];
  if (J.depth == 0) return;  // 1^3 grid: its leaf is the root, written with the job bases
  const uint64_t po = M.off[1] + leaf_parent[i];
  const uint32_t cnt = (uint32_t)__popc((uint32_t)A.mask[po]);
  const uint64_t cb = A.start[po] + A.size[po] - 1 - cnt;
  const uint64_t rank = i - A.first[po];
  const uint64_t pos = cb + rank;
  const uint64_t gpos = J.nbase + pos;
  if (pos >= J.n_nodes || gpos >= node_cap) {
    M.flag[1] = 1;
    return;
  }
  uint64_t data = J.payload ? (1ull + (i - J.voff)) : 1ull;
  if (A.isize && J.levels && J.payload) {
    const uint64_t idx = 1 + A.pstart[po] + rank;
    if (idx >= J.n_data) {
      M.flag[1] = 1;
      return;
    }
    const float4* src = (const float4*)(in_rec + (J.dbase + data) * 8);
    float4* dst = (float4*)(out_rec + (J.dbase + idx) * 8);
    dst[0] = src[0];
    dst[1] = src[1];
    data = idx;
  }
  put_node(J, nodes, gpos, data, 0ull, ~0ull);
}

// ---------------------------------------------------------------- brick fast path
// Every occupied 16^3 brick of the voxel stream is an aligned subtree of height 4 whose
// occupancy is its 64 words (word w = voxels 64w..64w+63 in Morton order). Within it:
//   height 1: a byte of a word (8 voxels), child mask = the byte;
//   height 2: a word, child mask = its non-zero bytes;
//   height 3: 8 consecutive words, child mask = its non-zero words;
//   height 4: the brick, child mask = its non-zero word groups.
// The voxelizer's brick kernels compute each brick node's subtree size and child mask
// (brick_shape_of, one wave per brick, lane = word) into the brick stream; here one wave
// per brick writes every record below the brick node at its post-order position, and
// the heights above come from the brick codes at once (upper levels, above).

__device__ __forceinline__ BrickShape brick_shape(const uint64_t* __restrict__ bwords,
                                                  uint64_t b) {
  return brick_shape_of(bwords[b * kBrickWords + lane_id()]);
}

// Every record below each brick node, written record by record: the brick's records form
// the contiguous range [S, S + s4 - 1) laid out as the word groups' subtree ranges, then
// the group nodes; a group's range is its words' ranges, then its word nodes; a word's
// range is its bytes' voxels, then its byte nodes. Per brick a wave builds the offsets of
// these ranges in LDS, then lane r writes records r, r + 64, ... of the range (looking
// its node up in the offsets), so consecutive lanes store consecutive records.
constexpr int kEmitWaves = 4;  // waves per block of k_emit_bricks

// last k in [0, 8) with a[k] <= v (a non-decreasing, a[0] <= v)
__device__ __forceinline__ int last_le8(const uint32_t* a, uint32_t v) {
  int k = 0;
  if (a[k + 4] <= v) k += 4;
  if (a[k + 2] <= v) k += 2;
  if (a[k + 1] <= v) k += 1;
  return k;
}

// kPlaced: the job's records go to a larger node stream (a streamed partition), their
// children_base and data indices moved by its placement.
template <bool kPlaced>
__global__ __launch_bounds__(kWave * kEmitWaves) void k_emit_bricks(
    Arena A, LevelMeta M, Tags T, const JobDev* __restrict__ jobs,
    const uint64_t* __restrict__ bwords, const uint64_t* __restrict__ bcode,
    const uint64_t* __restrict__ bvoff, uint64_t nb, uint64_t node_cap,
    uint64_t* __restrict__ nodes) {
  __shared__ uint32_t s_go[kEmitWaves][9];    // group ranges (relative to S)
  __shared__ uint32_t s_gs3[kEmitWaves][8];   // group subtree sizes
  __shared__ uint32_t s_gm3[kEmitWaves][8];   // group child masks (non-zero words)
  __shared__ uint32_t s_wo[kEmitWaves][64];   // word ranges (relative to S)
  __shared__ uint32_t s_s2[kEmitWaves][64];   // word subtree sizes
  __shared__ uint32_t s_m2[kEmitWaves][64];   // word child masks (non-zero bytes)
  __shared__ uint32_t s_vw[kEmitWaves][64];   // voxels before the word in the brick
  __shared__ uint64_t s_wd[kEmitWaves][64];   // the words
  __shared__ uint32_t s_bo[kEmitWaves][512];  // byte voxel ranges (relative to S)
  __shared__ uint64_t s_offs[256];            // offsets_of(m) for every child mask m
  __shared__ uint8_t s_nth[256][8];           // nth_bit(m, n)
  {
    const uint32_t m = threadIdx.x;  // the block has 256 threads: one mask each
    s_offs[m] = offsets_of(m);
    uint32_t x = m;
#pragma unroll
    for (int n = 0; n < 8; n++) {
      s_nth[m][n] = (uint8_t)(x ? __ffs(x) - 1 : 0);
      x &= x - 1;
    }
  }
  const int ws = threadIdx.x / kWave;
  const int lane = lane_id();
  const int g = lane >> 3, q = lane & 7;
  const uint64_t w = (uint64_t)blockIdx.x * kEmitWaves + ws;
  const bool live = w < nb && !M.flag[0];
  // every wave of the block reaches the barrier below; a wave without a brick works on
  // zeros and writes nothing
  // the brick's own words first, then its job and offsets: all in flight while the
  // offset tables are built
  uint64_t bc = 0, bvo = 0, S = 0;
  BrickShape B;
  if (live) {
    B = brick_shape(bwords, w);
    bc = bcode[w];
    bvo = bvoff[w];
    S = A.start[M.off[4] + w];
  } else {
    B.word = 0;
    B.m2 = B.s2 = B.mask3 = B.s3 = B.mask4 = 0;
    B.s4 = 1;
  }
  const uint32_t v3 = (q == 0 && B.s3) ? B.s3 - 1u : 0u;
  const uint32_t inc3 = wave_incl_scan<uint32_t>(v3);
  // only group heads add to inc3, and every lane of a group knows its group's s3: the
  // group's exclusive prefix is inc3 less the head's own term
  const uint32_t pre3 = inc3 - (B.s3 ? B.s3 - 1u : 0u);
  const uint32_t v2 = B.word ? B.s2 - 1u : 0u;
  // inclusive scan within each group of 8 lanes (DPP row shifts, lanes whose source lies
  // in the previous group add nothing)
  int x2 = (int)v2;
  {
    const int o1 = __builtin_amdgcn_update_dpp(0, x2, 0x111, 0xf, 0xf, true);  // row_shr:1
    x2 += (q >= 1) ? o1 : 0;
    const int o2 = __builtin_amdgcn_update_dpp(0, x2, 0x112, 0xf, 0xf, true);  // row_shr:2
    x2 += (q >= 2) ? o2 : 0;
    const int o4 = __builtin_amdgcn_update_dpp(0, x2, 0x114, 0xf, 0xf, true);  // row_shr:4
    x2 += (q >= 4) ? o4 : 0;
  }
  const uint32_t inc2 = (uint32_t)x2;
  const uint32_t wo = pre3 + (inc2 - v2);
  const uint32_t pw = (uint32_t)__popcll(B.word);
  const uint32_t incw = wave_incl_scan<uint32_t>(pw);
  if (q == 0) {
    s_go[ws][g] = pre3;
    s_gs3[ws][g] = B.s3;
    s_gm3[ws][g] = B.mask3;
  }
  if (lane == 0) s_go[ws][8] = B.s4 - 1u - (uint32_t)__popc(B.mask4);
  s_wo[ws][lane] = wo;
  s_s2[ws][lane] = B.s2;
  s_m2[ws][lane] = B.m2;
  s_vw[ws][lane] = incw - pw;
  s_wd[ws][lane] = B.word;
  uint32_t bo = wo;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    s_bo[ws][lane * 8 + j] = bo;
    bo += (uint32_t)__popc((uint32_t)(B.word >> (8 * j)) & 0xFFu);
  }
  __syncthreads();
  if (!live) return;
  const JobDev& J = jobs[T.job(bc)];
  const uint64_t vox0 = bvo - J.voff;
  const uint32_t R = B.s4 - 1u;        // records below the brick node
  const uint32_t RG = s_go[ws][8];     // records inside the groups' ranges
  if (S + R > J.n_nodes || J.nbase + S + R > node_cap) {
    if (lane == 0) M.flag[1] = 1;
    return;
  }
  uint64_t* out = nodes + 3 * (J.nbase + S);
  const int payload = J.payload;
  // the job's placement, read once (the record stores below could alias the job table)
  const uint64_t cb_add = kPlaced ? J.cb_add : 0, data_add = kPlaced ? J.data_add : 0;
  const uint64_t d_hdr = payload ? 1ull : 2ull;
  for (uint32_t r = lane; r < R; r += kWave) {
    // find the record's node: its child mask m (internal nodes) and the first two words
    uint64_t d0 = 0, d1;
    uint32_t m;
    bool leaf = false;
    if (r >= RG) {  // a group node
      const int gg = s_nth[B.mask4][r - RG];
      m = s_gm3[ws][gg];
      d1 = S + s_go[ws][gg] + s_gs3[ws][gg] - 1u - (uint32_t)__popc(m);
    } else {
      const int gg = last_le8(s_go[ws], r);
      const uint32_t m3 = s_gm3[ws][gg];
      const uint32_t r1 = r - s_go[ws][gg];
      const uint32_t l3 = s_gs3[ws][gg] - 1u, nw = (uint32_t)__popc(m3);
      if (r1 >= l3 - nw) {  // a word node
        const int ww = gg * 8 + s_nth[m3][r1 - (l3 - nw)];
        m = s_m2[ws][ww];
        d1 = S + s_wo[ws][ww] + s_s2[ws][ww] - 1u - (uint32_t)__popc(m);
      } else {
        const int ww = gg * 8 + last_le8(&s_wo[ws][gg * 8], r);
        const uint32_t m2 = s_m2[ws][ww];
        const uint32_t r2 = r - s_wo[ws][ww];
        const uint32_t l2 = s_s2[ws][ww] - 1u, nbz = (uint32_t)__popc(m2);
        const uint64_t word = s_wd[ws][ww];
        if (r2 >= l2 - nbz) {  // a byte node: its voxels start at its own range start
          const int j = s_nth[m2][r2 - (l2 - nbz)];
          m = (uint32_t)(word >> (8 * j)) & 0xFFu;
          d1 = S + s_bo[ws][ww * 8 + j];
        } else {  // a voxel
          const int j = last_le8(&s_bo[ws][ww * 8], r);
          const uint32_t k = r - s_bo[ws][ww * 8 + j];
          const uint32_t before = (uint32_t)__popcll(word & ((1ull << (8 * j)) - 1ull));
          d0 = payload ? 1ull + vox0 + s_vw[ws][ww] + before + k : 1ull;
          d1 = 0;
          m = 0;
          leaf = true;
        }
      }
    }
    const uint64_t d2 = leaf ? ~0ull : s_offs[m];
    if constexpr (kPlaced) {
      d0 = d0 >= d_hdr ? d0 + data_add : d0;
      d1 = leaf ? d1 : d1 + cb_add;
    }
    out[3 * (uint64_t)r] = d0;
    out[3 * (uint64_t)r + 1] = d1;
    out[3 * (uint64_t)r + 2] = d2;
  }
}

// ---------------------------------------------------------------- upper levels at once
// Fast path: every height >= 5 straight from the sorted brick nodes (height 4) in four
// launches, instead of one scan per height. Brick i heads its ancestor at height l (is the
// first brick below it) iff l <= T(i): T(i) = the job's depth D for the first brick of a
// job, else 4 + (highest bit where its code differs from brick i-1's) / 3, capped at D.
// So the index of brick i's ancestor at height l is incl_l(i) - 1, incl_l(i) = the number
// of bricks j <= i with T(j) >= l. With S4(i) = the brick subtree sizes before brick i and
//   A_l(i) = S4(i) + sum_{m=5..l} (bricks before i with T >= m)
// (the nodes below height l+1 that precede brick i in the stream), a node P of height l
// whose bricks are [i0, i1) has size(P) = A_l(i1) - A_l(i0), and the sizes of P's
// preceding siblings sum to A_l(head of P) - A_l(head of its parent's first child).
// k_up_count counts per tile of 256 bricks, k_up_scan scans the counts per height (one
// block), k_up_apply writes every node's code, first child, parent index and A_l at its
// head brick (and A_l at every job's end), k_up_finish turns them into sizes, child masks
// and sibling offsets. No atomics.
constexpr int kUpTile = 256;
constexpr int kUpLevels = 17;  // heights 5 .. 21

// T(i) from brick i's node code c and the previous brick's p (has_prev = i > 0)
__device__ __forceinline__ int brick_top(uint64_t c, uint64_t p, bool has_prev, Tags T, int D) {
  if (D <= 4) return 4;
  if (!has_prev || (c & T.tag_mask) != (p & T.tag_mask)) return D;
  const uint64_t x = (c ^ p) & ~T.tag_mask;
  if (!x) return 4;
  return min(D, 4 + (63 - __clzll((long long)x)) / 3);
}
// the height-4 node code of a brick (its first voxel's code without the 12 in-brick bits)
__device__ __forceinline__ uint64_t brick_node_code(uint64_t c0, Tags T) {
  return (c0 & T.tag_mask) | ((c0 & ~T.tag_mask) >> 12);
}
__device__ __forceinline__ int job_depth(const JobDev* jobs, Tags T, uint64_t code) {
  return T.tag_mask ? jobs[T.job(code)].depth : jobs[0].depth;
}

// Per brick: its height-4 node (code, subtree size and child mask from the voxelizer's
// brick shape) and the hazard case (the job's last voxel is code G^3 - 2: bits 62 and 63
// of the last word of the grid's last brick); per tile: counts of T >= l for l = 5 .. H,
// and the sum of the brick subtree sizes.
__global__ __launch_bounds__(kUpTile) void k_up_count(Arena A, Tags T, const JobDev* jobs,
                                                      const uint64_t* __restrict__ bwords,
                                                      const uint64_t* __restrict__ bcode,
                                                      const uint32_t* __restrict__ bshape,
                                                      uint64_t NB, int H,
                                                      uint32_t* __restrict__ tile_cnt,
                                                      uint64_t* __restrict__ tile_s4,
                                                      unsigned long long* hazard) {
  __shared__ uint32_t wcnt[kUpTile / kWave][kUpLevels];
  __shared__ uint64_t ws4[kUpTile / kWave];
  const uint64_t i = (uint64_t)blockIdx.x * kUpTile + threadIdx.x;
  int t = 4;
  uint64_t s4 = 0;
  if (i < NB) {
    const uint64_t c0 = bcode[i];
    const uint64_t code = brick_node_code(c0, T);
    const uint64_t prev = i ? brick_node_code(bcode[i - 1], T) : 0ull;
    const uint32_t sh = bshape[i];
    s4 = sh & 0xFFFFu;
    A.code[i] = code;
    A.size[i] = s4;
    A.mask[i] = (uint8_t)(sh >> 16);
    A.first[i] = 0;
    const uint32_t jb = T.job(c0);
    const JobDev& J = jobs[jb];
    const uint64_t G = J.gridsize;
    if ((c0 & ~T.tag_mask) == G * G * G - kBrickVox &&
        (bwords[i * kBrickWords + kWave - 1] >> 62) == 1ull)
      hazard[jb] = 1;
    t = brick_top(code, prev, i > 0, T, J.depth);
  }
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < kUpLevels; k++) {
    const uint64_t b = __ballot(t >= 5 + k);
    if (lane_id() == 0) wcnt[w][k] = (uint32_t)__popcll(b);
  }
  const uint64_t sw = wave_last<uint64_t>(wave_incl_scan<uint64_t>(s4));
  if (lane_id() == 0) ws4[w] = sw;
  __syncthreads();
  if (threadIdx.x < kUpLevels && 5 + (int)threadIdx.x <= H) {
    uint32_t sum = 0;
    for (int q = 0; q < kUpTile / kWave; q++) sum += wcnt[q][threadIdx.x];
    tile_cnt[(uint64_t)blockIdx.x * kUpLevels + threadIdx.x] = sum;
  }
  if (threadIdx.x == 0) {
    uint64_t sum = 0;
    for (int q = 0; q < kUpTile / kWave; q++) sum += ws4[q];
    tile_s4[blockIdx.x] = sum;
  }
}

// one block, a wave per height (and one for the sizes): the exclusive prefix of the tile
// counts (in place, 64 tiles at a time); then the node counts and arena offsets of all
// heights and A_l at the end of the stream; an arena too small is flagged
__global__ __launch_bounds__(1024) void k_up_scan(Arena A, LevelMeta M, uint64_t NB, int H,
                                                  uint32_t ntiles,
                                                  uint32_t* __restrict__ tile_cnt,
                                                  uint64_t* __restrict__ tile_s4,
                                                  uint64_t* __restrict__ a_end) {
  __shared__ uint64_t s_tot[kUpLevels + 1];
  const int w = threadIdx.x / kWave;
  for (int k = w; k <= kUpLevels; k += 1024 / kWave) {
    if (k < kUpLevels && 5 + k > H) continue;  // uniform per wave
    uint64_t carry = 0;
    for (uint32_t t0 = 0; t0 < ntiles; t0 += kWave) {
      const uint32_t t = t0 + lane_id();
      if (k == kUpLevels) {  // the brick subtree sizes
        const uint64_t v = t < ntiles ? tile_s4[t] : 0ull;
        const uint64_t inc = wave_incl_scan<uint64_t>(v);
        if (t < ntiles) tile_s4[t] = carry + inc - v;
        carry += wave_last<uint64_t>(inc);
      } else {
        uint32_t* cell = tile_cnt + (uint64_t)t * kUpLevels + k;
        const uint32_t v = t < ntiles ? *cell : 0u;
        const uint32_t inc = wave_incl_scan<uint32_t>(v);
        if (t < ntiles) *cell = (uint32_t)carry + inc - v;
        carry += wave_last<uint32_t>(inc);
      }
    }
    if (lane_id() == 0) s_tot[k] = carry;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t off = NB;  // off[5]
    uint64_t a = s_tot[kUpLevels];  // S4 at the end
    for (int l = 5; l <= H; l++) {
      const uint64_t n = s_tot[l - 5];
      M.n[l] = n;
      M.off[l] = off;
      off += n;
      M.off[l + 1] = off;
      a += n;
      a_end[l - 5] = a;
    }
    if (off > A.cap) M.flag[0] = 1;
  }
}

// Per brick: the parent index of its brick node, the fields of every node it heads, A_l at
// each of those nodes (kept in the node's start slot until the records are written), and
// at a job's first brick A_l as the previous job's end (the last brick: its job's end is
// the end of the stream).
__global__ __launch_bounds__(kUpTile) void k_up_apply(Arena A, LevelMeta M, Tags T,
                                                      const JobDev* jobs, uint64_t NB, int H,
                                                      const uint32_t* __restrict__ tile_pre,
                                                      const uint64_t* __restrict__ tile_s4,
                                                      const uint64_t* __restrict__ a_end,
                                                      uint64_t* __restrict__ job_end) {
  __shared__ uint32_t wpre[kUpTile / kWave][kUpLevels];
  __shared__ uint64_t ws4[kUpTile / kWave];
  if (M.flag[0]) return;  // arena too small: the build is redone with a larger one
  const uint64_t i = (uint64_t)blockIdx.x * kUpTile + threadIdx.x;
  const bool live = i < NB;
  uint64_t c = 0, s4 = 0, cp = 0;
  int D = 4, t = 4;
  if (live) {
    c = A.code[i];
    s4 = A.size[i];
    if (i) cp = A.code[i - 1];
    D = job_depth(jobs, T, c);
    t = brick_top(c, cp, i > 0, T, D);
  }
  const int w = threadIdx.x / kWave;
  const uint64_t le = (2ull << lane_id()) - 1ull;  // lanes <= this one
  uint64_t bal[kUpLevels];
#pragma unroll
  for (int k = 0; k < kUpLevels; k++) {
    bal[k] = __ballot(t >= 5 + k);
    if (lane_id() == 0) wpre[w][k] = (uint32_t)__popcll(bal[k]);
  }
  const uint64_t s4i = wave_incl_scan<uint64_t>(s4);
  if (lane_id() == kWave - 1) ws4[w] = s4i;
  __syncthreads();
  if (threadIdx.x < kUpLevels) {  // exclusive prefix over the waves, per height
    uint32_t run = 0;
    for (int q = 0; q < kUpTile / kWave; q++) {
      const uint32_t x = wpre[q][threadIdx.x];
      wpre[q][threadIdx.x] = run;
      run += x;
    }
  }
  __syncthreads();
  uint64_t S4 = tile_s4[blockIdx.x] + s4i - s4;
  for (int q = 0; q < w; q++) S4 += ws4[q];
  if (!live) return;
  const uint64_t tag = c & T.tag_mask, rest = c & ~T.tag_mask;
  const bool job_first = i == 0 || (cp & T.tag_mask) != tag;
  const uint32_t jb = T.job(c);
  A.start[i] = S4;  // A_4 at the brick
  if (D >= 5)
    A.pidx[i] = tile_pre[(uint64_t)blockIdx.x * kUpLevels] + wpre[w][0] +
                (uint32_t)__popcll(bal[0] & le) - 1u;
  uint64_t a = S4;
  uint32_t prev_incl = 0;
#pragma unroll
  for (int k = 0; k < kUpLevels; k++) {
    const int l = 5 + k;
    if (l > H) break;
    const uint32_t incl = tile_pre[(uint64_t)blockIdx.x * kUpLevels + k] + wpre[w][k] +
                          (uint32_t)__popcll(bal[k] & le);
    const bool head = t >= l;
    a += incl - (head ? 1u : 0u);  // A_l(i)
    if (head && l <= D) {
      const uint64_t g = M.off[l] + incl - 1u;
      A.code[g] = tag | (rest >> (3 * (l - 4)));
      A.first[g] = l == 5 ? (uint32_t)i : prev_incl - 1u;
      if (k + 1 < kUpLevels && l < D) {
        const uint32_t nxt = tile_pre[(uint64_t)blockIdx.x * kUpLevels + k + 1] +
                             wpre[w][k + 1] + (uint32_t)__popcll(bal[k + 1] & le);
        A.pidx[g] = nxt - 1u;
      }
      A.start[g] = a;
    }
    if (job_first && i) job_end[(uint64_t)T.job(cp) * kUpLevels + k] = a;
    if (i == NB - 1) job_end[(uint64_t)jb * kUpLevels + k] = a_end[k];
    prev_incl = incl;
  }
}

// Every node of heights 4 .. H: size and child mask (heights >= 5), sibling offset (below
// its job's root); each job: its tree size.
__global__ __launch_bounds__(256) void k_up_finish(Arena A, LevelMeta M, Tags T,
                                                   const JobDev* jobs, int H,
                                                   const uint64_t* __restrict__ job_end,
                                                   unsigned long long* root_size) {
  if (M.flag[0]) return;
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int l = 4;
  while (l <= H && g >= M.off[l] + M.n[l]) l++;
  if (l > H) return;
  const uint64_t co = M.off[l], i = g - co;
  const uint64_t code = A.code[g];
  const uint32_t jb = T.job(code);
  const int D = job_depth(jobs, T, code);
  const uint64_t a0 = A.start[g];
  if (l >= 5) {
    const bool last = i + 1 == M.n[l] || T.job(A.code[g + 1]) != jb;
    const uint64_t a1 = last ? job_end[(uint64_t)jb * kUpLevels + (l - 5)] : A.start[g + 1];
    const uint64_t size = a1 - a0;
    A.size[g] = size;
    // children: the nodes of height l-1 from the first child on whose parent is this node
    const uint64_t cco = M.off[l - 1], nch = M.n[l - 1];
    const uint64_t f = A.first[g];
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (f + k < nch) {
        const uint64_t cc = A.code[cco + f + k];
        if (T.parent(cc) == code) m |= 1u << (cc & 7);
      }
    }
    A.mask[g] = (uint8_t)m;
    if (l == D) root_size[jb] = size;
  }
  if (l == 4 && D == 4) root_size[jb] = A.size[g];  // a 16^3 job: the brick is the root
  if (l < D) {
    const uint64_t p = M.off[l + 1] + A.pidx[g];
    const uint64_t f = A.first[p];
    A.sib[g] = a0 - A.start[co + f] - (i - f);
  }
}

#define SVO_TRY(x)                   \
  do {                               \
    hipError_t _e = (x);             \
    if (_e != hipSuccess) return _e; \
  } while (0)

hipError_t build_jobs(Context& c, uint32_t n_jobs, bool tagged, int levels,
                      const BuildOut* out) {
  if ((levels || out) && (tagged || n_jobs != 1)) {
    set_error(levels ? "-levels is supported for single builds only"
                     : "internal: a placed build is a single build");
    return hipErrorNotSupported;
  }
  const uint64_t V = c.n_vox;
  if (V >= (1ull << 32)) {
    set_error("more than 2^32 voxels in one build");
    return hipErrorInvalidValue;
  }
  if (n_jobs == 0 || n_jobs > kMaxJobs || (!tagged && n_jobs != 1)) {
    set_error("invalid job count");
    return hipErrorInvalidValue;
  }
  // host view of the jobs: depth, voxel range, data range
  std::vector<JobDev> jd(n_jobs);
  uint64_t n_records = 0;
  int H = 0;
  std::vector<uint64_t> bound;
  for (uint32_t j = 0; j < n_jobs; j++) {
    Job& J = c.jobs[j];
    JobDev& d = jd[j];
    memset(&d, 0, sizeof(d));
    int D = 0;
    while ((1ull << D) < J.gridsize) D++;
    if (tagged && D > 16) {
      set_error("a batched job's gridsize is at most 65536");
      return hipErrorInvalidValue;
    }
    d.tris = J.tris;
    d.unit_div = J.unit_vox != 0.0f ? 1.0f / J.unit_vox : 0.0f;
    d.color_mode = J.color_mode;
    d.payload = J.payload;
    d.depth = D;
    d.levels = levels ? 1 : 0;
    d.gridsize = J.gridsize;
    d.voff = J.voff;
    d.nvox = J.nvox;
    J.dbase = n_records;
    J.n_data = J.payload ? 1 + J.nvox : 2;
    d.dbase = J.dbase;
    if (out) {
      d.cb_add = out->cb_add;
      d.data_add = out->data_add;
    }
    n_records += J.n_data;
    if (J.nvox) H = std::max(H, D);
  }
  // brick fast path: the four lowest levels come from the voxelizer's brick stream (not
  // for -levels, for voxels added from outside, or for meshes below 16^3)
  bool fast = c.bricks_ok && !levels && c.n_bricks > 0;
  for (uint32_t j = 0; j < n_jobs; j++)
    if (jd[j].nvox && jd[j].depth < 4) fast = false;
  if (!fast) SVO_TRY(ensure_codes(c));  // the level scans start from the codes
  const uint64_t NB = fast ? c.n_bricks : 0;
  const int h0 = fast ? 4 : 0;  // lowest height the level scans start from
  // per-height upper bounds (host) size the launch grids
  bound.assign(H + 2, 0);
  bound[0] = V;
  for (int h = 1; h <= H; h++) {
    uint64_t b = 0;
    for (uint32_t j = 0; j < n_jobs; j++) {
      if (!jd[j].nvox || jd[j].depth < h) continue;
      const int e = jd[j].depth - h;
      b += (3 * e >= 63) ? jd[j].nvox : std::min<uint64_t>(jd[j].nvox, 1ull << (3 * e));
    }
    bound[h] = std::min(bound[h - 1], b);
    if (fast && h == 4) bound[h] = NB;
  }
  // the voxelizer's last upload usually holds these already
  if (c.jobdev_img.size() != jd.size() ||
      memcmp(c.jobdev_img.data(), jd.data(), n_jobs * sizeof(JobDev)) != 0) {
    SVO_TRY(ensure(c.jobdev, n_jobs * sizeof(JobDev)));
    SVO_TRY(upload_async(c, c.jobdev.p, jd.data(), n_jobs * sizeof(JobDev)));
  }
  c.jobdev_img.clear();  // the job bases are written into it
  // data records: payload records were written by the voxelizer at dbase + 1 ..
  SVO_TRY(ensure_keep(c.records, n_records * 32, n_records * 32, c.stream));
  JobDev* d_jobs = c.jobdev.as<JobDev>();
  const Tags T{tagged ? ~kCodeMask : 0ull};

  // arena: the internal nodes the level scans produce (heights >= h0)
  uint64_t cap = fast ? NB + NB / 2 + 4096 : std::max<uint64_t>(V / 2 + 1024, 4096);
  // internal nodes below the brick nodes (fast path): bytes, words, word groups
  const uint64_t in_brick =
      fast ? std::min<uint64_t>(V, 512 * NB) + std::min<uint64_t>(V, 64 * NB) + 8 * NB : 0;
  for (int attempt = 0;; attempt++) {
    SVO_TRY(ensure(c.lvl_code, cap * 8));
    SVO_TRY(ensure(c.lvl_size, cap * 8));
    SVO_TRY(ensure(c.lvl_mask, cap + 8));  // (+: masks are or-ed as 32-bit words)
    SVO_TRY(ensure(c.lvl_first, cap * 4));
    SVO_TRY(ensure(c.lvl_start, cap * 8));
    SVO_TRY(ensure(c.lvl_pidx, cap * 4));
    SVO_TRY(ensure(c.lvl_sib, cap * 8));
    // meta | per job: root size, root internal size, hazard | single-pass scan status
    const uint64_t meta_words = 2 * 64 + 8 + 3 * (uint64_t)n_jobs;
    // heights from h_tail up go to the single-block tail kernel; on the fast path all
    // heights above the bricks come from the brick codes at once (no level scans)
    const bool up_once = fast;
    int h_tail = H;
    for (int h = h0; h < H && !up_once; h++)
      if (bound[h] <= kTailLevelMax) {
        h_tail = h;
        break;
      }
    const int h_scan_end = up_once ? h0 : h_tail;  // level scans for h0 .. h_scan_end-1
    uint64_t st_words = 0;
    for (int h = h0; h < h_scan_end; h++) st_words += scan_single_words(bound[h]);
    SVO_TRY(ensure(c.lvl_aux, (meta_words + st_words) * 8));
    if (levels) {
      SVO_TRY(ensure(c.lvl_isize, cap * 8));
      SVO_TRY(ensure(c.lvl_cache, cap * 24));
      SVO_TRY(ensure(c.lvl_pstart, cap * 8));
      // null (+ white) record, one per voxel (payload), one per internal node (<= cap)
      SVO_TRY(ensure(c.records2, (n_records + cap) * 32));
    }
    float* out_rec = levels ? c.records2.as<float>() : c.records.as<float>();
    if (!fast) SVO_TRY(ensure(c.lvl_leafp, std::max<uint64_t>(V, 1) * 4));
    const uint64_t node_cap = cap + V + in_brick + n_jobs;
    uint64_t* nodes;
    if (out) {
      SVO_TRY(ensure_keep(*out->nodes, (out->at + node_cap) * 24, out->at * 24, c.stream));
      nodes = out->nodes->as<uint64_t>() + 3 * out->at;
    } else {
      SVO_TRY(ensure(c.nodes, node_cap * 24));
      nodes = c.nodes.as<uint64_t>();
    }
    Arena A{c.lvl_code.as<uint64_t>(), c.lvl_size.as<uint64_t>(), c.lvl_mask.as<uint8_t>(),
            c.lvl_first.as<uint32_t>(), c.lvl_start.as<uint64_t>(), c.lvl_pidx.as<uint32_t>(),
            cap, levels ? c.lvl_isize.as<uint64_t>() : nullptr,
            levels ? c.lvl_cache.as<float>() : nullptr,
            levels ? c.lvl_pstart.as<uint64_t>() : nullptr, c.lvl_sib.as<uint64_t>()};
    unsigned long long* meta = c.lvl_aux.as<unsigned long long>();
    LevelMeta M{meta, meta + 64, meta + 128};
    unsigned long long* root_size = meta + 136;
    unsigned long long* root_isize = root_size + n_jobs;
    unsigned long long* brick_hazard = root_isize + n_jobs;
    // the upper levels' tile scratch (fast path)
    const uint32_t ntiles = (uint32_t)((NB + kUpTile - 1) / kUpTile);
    const uint64_t b_cnt = ((uint64_t)ntiles * kUpLevels * 4 + 255) & ~255ull;
    const uint64_t b_s4 = ((uint64_t)ntiles * 8 + 255) & ~255ull;
    const uint64_t b_end = ((uint64_t)kUpLevels * 8 + 255) & ~255ull;
    if (up_once)
      SVO_TRY(ensure(c.up_tiles, b_cnt + b_s4 + b_end + (uint64_t)n_jobs * kUpLevels * 8));
    // the build's kernels up to its read-back: one graph segment
    SVO_TRY(run_segment(c, [&]() -> hipError_t {
      // n[0] = V, off[1] = 0, flags = 0, root sizes = 0, hazard flags = 0, scan status = 0
      // (fast: n[4] = NB, off[4] = 0, off[5] = NB)
      SVO_LAUNCH(k_build_init, dim3(1), dim3(256), 0, c.stream, meta,
                         meta_words + st_words, V, NB, fast ? 1 : 0);
      SVO_TRY(hipGetLastError());

      if (up_once) {
        char* ub = (char*)c.up_tiles.p;
        uint32_t* tile_cnt = (uint32_t*)ub;
        uint64_t* tile_s4 = (uint64_t*)(ub + b_cnt);
        uint64_t* a_end = (uint64_t*)(ub + b_cnt + b_s4);
        uint64_t* job_end = (uint64_t*)(ub + b_cnt + b_s4 + b_end);
        ktimer_begin(c, "upper_levels");
        SVO_LAUNCH(k_up_count, dim3(ntiles), dim3(kUpTile), 0, c.stream, A, T,
                           (const JobDev*)d_jobs, c.bwords.as<uint64_t>(),
                           c.bcode.as<uint64_t>(), c.bshape.as<uint32_t>(), NB, H, tile_cnt,
                           tile_s4, brick_hazard);
        SVO_LAUNCH(k_up_scan, dim3(1), dim3(1024), 0, c.stream, A, M, NB, H, ntiles,
                           tile_cnt, tile_s4, a_end);
        SVO_LAUNCH(k_up_apply, dim3(ntiles), dim3(kUpTile), 0, c.stream, A, M, T,
                           (const JobDev*)d_jobs, NB, H, (const uint32_t*)tile_cnt,
                           (const uint64_t*)tile_s4, (const uint64_t*)a_end, job_end);
        uint64_t nup = 0;
        for (int h = 4; h <= H; h++) nup += bound[h];
        SVO_LAUNCH(k_up_finish, dim3((unsigned)((nup + 255) / 256)), dim3(256), 0,
                           c.stream, A, M, T, (const JobDev*)d_jobs, H,
                           (const uint64_t*)job_end, root_size);
        ktimer_end(c);
        SVO_TRY(hipGetLastError());
      }

      // ---- bottom-up: heights 1 .. H
      unsigned long long* status = meta + meta_words;
      for (int h = h0; h < h_scan_end; h++) {
        HeadLoad ld{A, M, T, d_jobs, c.codes.as<uint64_t>(), h};
        ParentStore st{A, M, T, d_jobs, c.codes.as<uint64_t>(), c.lvl_leafp.as<uint32_t>(),
                       root_size, root_isize, c.records.as<float>() + 8, h};
        SVO_TRY((scan_single<uint32_t>(ld, st, bound[h], status,
                                       c.counters.as<unsigned long long>() + kCtrError, c.stream,
                                       M.n + h)));
        status += scan_single_words(bound[h]);
      }
      SVO_LAUNCH(k_levels_tail, dim3(1), dim3(kTailBlock), 0, c.stream, A, M, T, d_jobs,
                         n_jobs, c.codes.as<uint64_t>(), c.lvl_leafp.as<uint32_t>(), root_size,
                         root_isize, c.records.as<float>() + 8, h_tail, H,
                         fast ? brick_hazard : nullptr, node_cap,
                         nodes, c.records.as<float>(), out_rec);
      SVO_TRY(hipGetLastError());
      // ---- top-down: heights H .. 1 (one launch without -levels), then the leaves
      const int hs = std::max(h0, 1);
      if (!levels && H >= hs) {
        uint64_t nup = 0;
        for (int h = hs; h <= H; h++) nup += bound[h];
        ktimer_begin(c, "emit_children");
        SVO_LAUNCH(k_emit_upper, dim3((unsigned)std::max<uint64_t>((nup + 255) / 256, 1)),
                           dim3(256), 0, c.stream, A, M, T, (const JobDev*)d_jobs, hs, H, node_cap,
                           nodes);
        ktimer_end(c);
        SVO_TRY(hipGetLastError());
      }
      for (int h = H; levels && h >= hs; h--) {
        ktimer_begin(c, "emit_children");
        SVO_LAUNCH(k_emit_level, dim3((unsigned)std::max<uint64_t>((bound[h] + 255) / 256, 1)),
                           dim3(256), 0, c.stream, A, M, T, (const JobDev*)d_jobs, h, node_cap,
                           nodes, out_rec);
        ktimer_end(c);
        SVO_TRY(hipGetLastError());
      }
      if (fast) {
        ktimer_begin(c, "emit_bricks");
        const dim3 eg((unsigned)((NB + kEmitWaves - 1) / kEmitWaves)), eb(kWave * kEmitWaves);
        if (out)
          SVO_LAUNCH(k_emit_bricks<true>, eg, eb, 0, c.stream, A, M, T, (const JobDev*)d_jobs,
                     c.bwords.as<uint64_t>(), c.bcode.as<uint64_t>(), c.bvoff.as<uint64_t>(),
                     NB, node_cap, nodes);
        else
          SVO_LAUNCH(k_emit_bricks<false>, eg, eb, 0, c.stream, A, M, T, (const JobDev*)d_jobs,
                     c.bwords.as<uint64_t>(), c.bcode.as<uint64_t>(), c.bvoff.as<uint64_t>(),
                     NB, node_cap, nodes);
        ktimer_end(c);
        SVO_TRY(hipGetLastError());
      } else if (V) {
        ktimer_begin(c, "emit_leaves");
        SVO_LAUNCH(k_emit_leaves, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, c.stream,
                           A, M, T, (const JobDev*)d_jobs, node_cap, c.codes.as<uint64_t>(),
                           c.lvl_leafp.as<uint32_t>(), nodes,
                           c.records.as<float>(), out_rec);
        ktimer_end(c);
        SVO_TRY(hipGetLastError());
      }
      return hipSuccess;
    }));
    unsigned long long flags[3] = {0, 0, 0};
    if (c.build_done_ev) SVO_TRY(hipEventRecord(c.build_done_ev, c.stream));
    {
      const ReadReq rq[3] = {{flags, M.flag, 24},
                             {jd.data(), d_jobs, n_jobs * sizeof(JobDev)},
                             {&c.err_count, c.counters.as<unsigned long long>() + kCtrError, 8}};
      SVO_TRY(read_back(c, rq, 3));
    }
    ktimer_flush(c);
    if (flags[1] && !flags[0]) {
      set_error("internal: node record out of range");
      return hipErrorUnknown;
    }
    if (!flags[0]) {
      c.n_nodes = flags[2];
      break;
    }
    if (attempt > 8) {
      set_error("internal: level arena could not be sized");
      return hipErrorUnknown;
    }
    cap *= 4;
  }
  n_records = 0;
  for (uint32_t j = 0; j < n_jobs; j++) {
    Job& J = c.jobs[j];
    J.n_nodes = jd[j].n_nodes;
    J.nbase = jd[j].nbase;
    J.hazard = (int)jd[j].hazard;
    J.n_data = jd[j].n_data;
    n_records += J.n_data;
  }
  c.n_data = n_records;
  c.data_in_records2 = levels != 0;  // the post-order records are the output
  return hipSuccess;
}

}  // namespace svo
