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
// one scan per level; top-down passes give every parent thread its <= 8 children's
// starts and write their 24-byte records (Node.h:16-18, octree_io.h:62-66).
#include <algorithm>
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

__device__ __forceinline__ void put_node(uint64_t* nodes, uint64_t pos, uint64_t data,
                                         uint64_t cb, uint64_t offs) {
  uint64_t* r = nodes + 3 * pos;
  r[0] = data;
  r[1] = cb;
  r[2] = offs;
}

// Level arrays live in arenas (code, size, mask, first child, start), level l at element
// offset off[l]; the per-level counts n[l] and offsets are kept on the device so that the
// passes run back to back without host round trips. Level D (the leaves) is the voxel
// code stream itself, with implicit size 1.
struct LevelMeta {
  unsigned long long* n;    // n[l], l = 0..D
  unsigned long long* off;  // off[l], l = 0..D-1
  unsigned long long* flag; // [0] arena overflow, [1] node range error, [2] root size
};

struct Arena {
  uint64_t* code;
  uint64_t* size;
  uint8_t* mask;
  uint32_t* first;
  uint64_t* start;
  uint64_t cap;
};

// Parent level l from child level l+1: head = first child of a parent group.
struct HeadLoad {
  Arena A;
  LevelMeta M;
  const uint64_t* leaves;
  int l, D;
  __device__ uint32_t operator()(uint64_t i) const {
    // an overflowed child level is not read (its tail lies beyond the arena)
    if (l + 1 < D && M.off[l + 1] + M.n[l + 1] > A.cap) return 0u;
    const uint64_t* code = (l + 1 == D) ? leaves : A.code + M.off[l + 1];
    return (i == 0 || (code[i] >> 3) != (code[i - 1] >> 3)) ? 1u : 0u;
  }
};
struct ParentStore {
  Arena A;
  LevelMeta M;
  const uint64_t* leaves;
  uint32_t* leaf_parent;  // leaf level only: parent index of every leaf
  int l, D;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t head) const {
    const bool leaf = (l + 1 == D);
    if (leaf) leaf_parent[i] = ex + head - 1;
    if (!leaf && M.off[l + 1] + M.n[l + 1] > A.cap) {
      M.flag[0] = 1;
      return;
    }
    const uint64_t* code = leaf ? leaves : A.code + M.off[l + 1];
    const uint64_t* csize = leaf ? nullptr : A.size + M.off[l + 1];
    const uint64_t nch = M.n[l + 1];
    const uint64_t base = M.off[l];
    if (head) {
      if (base + ex >= A.cap) {
        M.flag[0] = 1;
      } else {
        const uint64_t pc = code[i] >> 3;
        uint32_t m = 0;
        uint64_t sz = 1;
        for (uint64_t j = i; j < nch && j < i + 8; j++) {
          const uint64_t cj = code[j];
          if ((cj >> 3) != pc) break;
          m |= 1u << (cj & 7);
          sz += csize ? csize[j] : 1ull;
        }
        A.code[base + ex] = pc;
        A.size[base + ex] = sz;
        A.mask[base + ex] = (uint8_t)m;
        A.first[base + ex] = (uint32_t)i;
      }
    }
    if (i == nch - 1) {
      const uint64_t np = (uint64_t)ex + head;
      M.n[l] = np;
      if (l > 0) M.off[l - 1] = base + np;
    }
  }
};

// One thread per parent at level l writes its children's records (level l+1).
__global__ __launch_bounds__(256) void k_emit_children(Arena A, LevelMeta M, int l, int D,
                                                       int payload, uint64_t node_cap,
                                                       uint64_t* __restrict__ nodes) {
  if (M.flag[0]) return;  // arena overflow: the build is redone with a larger arena
  const uint64_t np = M.n[l];
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const bool child_is_leaf = (l + 1 == D);
  const uint64_t po = M.off[l] + p;
  const uint64_t co = child_is_leaf ? 0 : M.off[l + 1];
  const uint32_t m = A.mask[po];
  const uint32_t cnt = (uint32_t)__popc(m);
  const uint64_t st = A.start[po];
  const uint64_t cb = st + A.size[po] - 1 - cnt;
  const uint32_t f = A.first[po];
  if (cb + cnt > node_cap) {
    M.flag[1] = 1;
    return;
  }
  if (child_is_leaf) return;  // leaves are written by k_emit_leaves
  uint64_t run = st;
  for (uint32_t r = 0; r < cnt; r++) {
    const uint64_t c = (uint64_t)f + r;
    const uint64_t sz = A.size[co + c];
    const uint32_t cm = A.mask[co + c];
    A.start[co + c] = run;
    put_node(nodes, cb + r, 0ull, run + sz - 1 - (uint32_t)__popc(cm), offsets_of(cm));
    run += sz - 1;
  }
}

// One thread per leaf: position = children_base(parent) + rank among its siblings.
// Consecutive leaves land in consecutive records, so the stores coalesce.
// OctreeBuilder.cpp:137-139 (binary: data 1) / :157-159 (payload: its data index).
__global__ __launch_bounds__(256) void k_emit_leaves(Arena A, LevelMeta M, int D, int payload,
                                                     uint64_t node_cap,
                                                     const uint32_t* __restrict__ leaf_parent,
                                                     uint64_t* __restrict__ nodes) {
  if (M.flag[0]) return;
  const uint64_t nl = M.n[D];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl) return;
  const uint64_t po = M.off[D - 1] + leaf_parent[i];
  const uint32_t cnt = (uint32_t)__popc((uint32_t)A.mask[po]);
  const uint64_t cb = A.start[po] + A.size[po] - 1 - cnt;
  const uint64_t pos = cb + (i - A.first[po]);
  if (pos >= node_cap) {
    M.flag[1] = 1;
    return;
  }
  put_node(nodes, pos, payload ? (1ull + i) : 1ull, 0ull, ~0ull);
}

__global__ void k_emit_root(Arena A, LevelMeta M, uint64_t node_cap, uint64_t* nodes) {
  if (M.flag[0] || M.n[0] != 1) {  // arena overflow, or not a single root
    M.flag[0] = 1;
    return;
  }
  const uint64_t o = M.off[0];
  const uint32_t m = A.mask[o];
  const uint64_t sz = A.size[o];
  A.start[o] = 0;
  M.flag[2] = sz;
  if (sz > node_cap) {
    M.flag[1] = 1;
    return;
  }
  put_node(nodes, sz - 1, 0ull, sz - 1 - (uint32_t)__popc(m), offsets_of(m));
}

#define SVO_TRY(x)                   \
  do {                               \
    hipError_t _e = (x);             \
    if (_e != hipSuccess) return _e; \
  } while (0)

hipError_t build_octree(Context& c, uint64_t gridsize, int levels) {
  if (levels) {
    set_error("-levels is not implemented on the GPU path");
    return hipErrorNotSupported;
  }
  int D = 0;
  while ((1ull << D) < gridsize) D++;
  const uint64_t V = c.n_vox;
  const uint64_t total_codes = gridsize * gridsize * gridsize;
  c.hazard = 0;
  if (V >= (1ull << 32)) {
    set_error("more than 2^32 voxels in one octree");
    return hipErrorInvalidValue;
  }

  // data records (octree_io.h:49-53, OctreeBuilder.cpp:25-29)
  if (!c.payload) {
    static const float white[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0};
    SVO_TRY(ensure(c.data, 64));
    SVO_TRY(hipMemcpyAsync(c.data.p, white, 64, hipMemcpyHostToDevice, c.stream));
    c.n_data = 2;
  } else {
    SVO_TRY(ensure_keep(c.records, (1 + V) * 32, (1 + V) * 32, c.stream));
    SVO_TRY(fill_u64(c.records.as<uint64_t>(), 0, 4, c.stream));
    c.n_data = 1 + V;
  }

  if (V == 0 || D == 0) {
    // empty grid: one null root (finalizeTree pads the whole grid); 1^3 grid with its
    // single voxel: the root is that leaf
    if (V == 0 && gridsize == 1) c.hazard = 1;
    const uint64_t rec[3] = {V ? 1ull : 0ull, 0, ~0ull};
    SVO_TRY(ensure(c.nodes, 24));
    SVO_TRY(hipMemcpyAsync(c.nodes.p, rec, 24, hipMemcpyHostToDevice, c.stream));
    c.n_nodes = 1;
    return hipStreamSynchronize(c.stream);
  }

  // per-level upper bounds (host) size the launch grids
  std::vector<uint64_t> bound(D + 1);
  for (int l = 0; l <= D; l++) bound[l] = (3 * l >= 63) ? V : std::min<uint64_t>(V, 1ull << (3 * l));
  uint64_t cap = std::max<uint64_t>(V / 2 + 1024, 4096);
  for (int attempt = 0;; attempt++) {
    SVO_TRY(ensure(c.lvl_code, cap * 8));
    SVO_TRY(ensure(c.lvl_size, cap * 8));
    SVO_TRY(ensure(c.lvl_mask, cap));
    SVO_TRY(ensure(c.lvl_first, cap * 4));
    SVO_TRY(ensure(c.lvl_start, cap * 8));
    SVO_TRY(ensure(c.lvl_aux, (2 * 64 + 8) * 8));
    SVO_TRY(ensure(c.lvl_leafp, V * 4));
    const uint64_t node_cap = cap + V + 1;
    SVO_TRY(ensure(c.nodes, node_cap * 24));
    SVO_TRY(ensure(c.scan_scratch, scan_scratch_elems(V) * 4 + 256));
    Arena A{c.lvl_code.as<uint64_t>(), c.lvl_size.as<uint64_t>(), c.lvl_mask.as<uint8_t>(),
            c.lvl_first.as<uint32_t>(), c.lvl_start.as<uint64_t>(), cap};
    unsigned long long* meta = c.lvl_aux.as<unsigned long long>();
    LevelMeta M{meta, meta + 64, meta + 128};
    // n[D] = V, off[D-1] = 0, flags = 0
    SVO_TRY(fill_u64((uint64_t*)meta, 0, 2 * 64 + 8, c.stream));
    SVO_TRY(hipMemcpyAsync(meta + D, &V, 8, hipMemcpyHostToDevice, c.stream));

    // ---- bottom-up: levels D-1 .. 0
    for (int l = D - 1; l >= 0; l--) {
      HeadLoad ld{A, M, c.codes.as<uint64_t>(), l, D};
      ParentStore st{A, M, c.codes.as<uint64_t>(), c.lvl_leafp.as<uint32_t>(), l, D};
      SVO_TRY((scan_exclusive<uint32_t>(ld, st, bound[l + 1], c.scan_scratch.as<uint32_t>(),
                                        c.stream, M.n + (l + 1))));
    }
    // ---- top-down
    hipLaunchKernelGGL(k_emit_root, dim3(1), dim3(1), 0, c.stream, A, M, node_cap,
                       c.nodes.as<uint64_t>());
    SVO_TRY(hipGetLastError());
    for (int l = 0; l + 1 < D; l++) {
      ktimer_begin(c, "emit_children");
      hipLaunchKernelGGL(k_emit_children, dim3((unsigned)((bound[l] + 255) / 256)), dim3(256), 0,
                         c.stream, A, M, l, D, c.payload, node_cap, c.nodes.as<uint64_t>());
      ktimer_end(c);
      SVO_TRY(hipGetLastError());
    }
    ktimer_begin(c, "emit_leaves");
    hipLaunchKernelGGL(k_emit_leaves, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, c.stream, A,
                       M, D, c.payload, node_cap, c.lvl_leafp.as<uint32_t>(),
                       c.nodes.as<uint64_t>());
    ktimer_end(c);
    SVO_TRY(hipGetLastError());
    unsigned long long flags[3] = {0, 0, 0};
    SVO_TRY(hipMemcpyAsync(flags, M.flag, 24, hipMemcpyDeviceToHost, c.stream));
    SVO_TRY(hipStreamSynchronize(c.stream));
    ktimer_flush(c);
    if (flags[1]) {
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
  {
    uint64_t last = 0;
    SVO_TRY(hipMemcpy(&last, c.codes.as<uint64_t>() + V - 1, 8, hipMemcpyDeviceToHost));
    // OctreeBuilder.cpp:36-42 leaves the root undefined when the last voxel is the
    // second-to-last code; the completed tree is built and the hazard reported
    if (total_codes >= 2 && last == total_codes - 2) c.hazard = 1;
  }
  return hipSuccess;
}

}  // namespace svo
