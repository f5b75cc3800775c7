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

struct LevelView {
  const uint64_t* code;  // sorted codes of this level
  const uint64_t* size;  // subtree sizes (nullptr on the leaf level: all 1)
  uint64_t n;
};

// Parent level from a child level: head = first child of a parent group.
struct HeadLoad {
  const uint64_t* code;
  __device__ uint32_t operator()(uint64_t i) const {
    return (i == 0 || (code[i] >> 3) != (code[i - 1] >> 3)) ? 1u : 0u;
  }
};
struct ParentStore {
  LevelView ch;
  uint64_t* pcode;
  uint64_t* psize;
  uint8_t* pmask;
  uint32_t* pfirst;
  ull* total;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t head) const {
    if (head) {
      const uint64_t pc = ch.code[i] >> 3;
      uint32_t m = 0;
      uint64_t sz = 1;
      for (uint64_t j = i; j < ch.n && j < i + 8; j++) {
        const uint64_t cj = ch.code[j];
        if ((cj >> 3) != pc) break;
        m |= 1u << (cj & 7);
        sz += ch.size ? ch.size[j] : 1ull;
      }
      pcode[ex] = pc;
      psize[ex] = sz;
      pmask[ex] = (uint8_t)m;
      pfirst[ex] = (uint32_t)i;
    }
    if (i == ch.n - 1) *total = (ull)(ex + head);
  }
};

// One thread per parent at level l writes its children's records (level l+1).
__global__ __launch_bounds__(256) void k_emit_children(
    uint64_t np, const uint64_t* __restrict__ pstart, const uint64_t* __restrict__ psize,
    const uint8_t* __restrict__ pmask, const uint32_t* __restrict__ pfirst,
    const uint64_t* __restrict__ csize, const uint8_t* __restrict__ cmask,
    uint64_t* __restrict__ cstart, int child_is_leaf, int payload, uint64_t leaf_base,
    uint64_t* __restrict__ nodes) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const uint32_t m = pmask[p];
  const uint32_t cnt = (uint32_t)__popc(m);
  const uint64_t st = pstart[p];
  const uint64_t cb = st + psize[p] - 1 - cnt;
  const uint32_t f = pfirst[p];
  uint64_t run = st;
  for (uint32_t r = 0; r < cnt; r++) {
    const uint64_t c = (uint64_t)f + r;
    if (child_is_leaf) {
      // OctreeBuilder.cpp:137-139 (binary: data 1) / :157-159 (payload: its data index)
      put_node(nodes, cb + r, payload ? (leaf_base + c) : 1ull, 0ull, ~0ull);
    } else {
      const uint64_t sz = csize[c];
      const uint32_t cm = cmask[c];
      cstart[c] = run;
      put_node(nodes, cb + r, 0ull, run + sz - 1 - (uint32_t)__popc(cm), offsets_of(cm));
      run += sz - 1;
    }
  }
}

__global__ void k_emit_root(const uint64_t* size, const uint8_t* mask, uint64_t* start,
                            uint64_t* nodes) {
  const uint32_t m = mask[0];
  const uint64_t sz = size[0];
  start[0] = 0;
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

  if (V == 0) {  // one null root (finalizeTree padding the whole grid)
    if (gridsize == 1) c.hazard = 1;
    static const uint64_t root[3] = {0, 0, ~0ull};
    SVO_TRY(ensure(c.nodes, 24));
    SVO_TRY(hipMemcpyAsync(c.nodes.p, root, 24, hipMemcpyHostToDevice, c.stream));
    c.n_nodes = 1;
    return hipStreamSynchronize(c.stream);
  }
  {
    uint64_t last = 0;
    SVO_TRY(hipMemcpyAsync(&last, c.codes.as<uint64_t>() + V - 1, 8, hipMemcpyDeviceToHost,
                           c.stream));
    SVO_TRY(hipStreamSynchronize(c.stream));
    // OctreeBuilder.cpp:36-42 leaves the root undefined when the last voxel is the
    // second-to-last code; the completed tree is built and the hazard reported
    if (total_codes >= 2 && last == total_codes - 2) c.hazard = 1;
  }
  if (D == 0) {  // 1^3 grid with its single voxel: the root is the leaf
    const uint64_t leaf[3] = {c.payload ? 1ull : 1ull, 0, ~0ull};
    SVO_TRY(ensure(c.nodes, 24));
    SVO_TRY(hipMemcpyAsync(c.nodes.p, leaf, 24, hipMemcpyHostToDevice, c.stream));
    c.n_nodes = 1;
    return hipStreamSynchronize(c.stream);
  }

  // ---- bottom-up: levels D-1 .. 0. Level arrays are packed one after another in
  // arenas; off[l] is the first element of level l.
  std::vector<uint64_t> n(D + 1, 0), off(D + 1, 0);
  n[D] = V;
  uint64_t cap = std::max<uint64_t>(V + V / 2 + 64, 1024), used = 0;
  auto alloc_levels = [&](uint64_t nc) -> hipError_t {
    SVO_TRY(ensure_keep(c.lvl_code, nc * 8, used * 8, c.stream));
    SVO_TRY(ensure_keep(c.lvl_size, nc * 8, used * 8, c.stream));
    SVO_TRY(ensure_keep(c.lvl_mask, nc, used, c.stream));
    SVO_TRY(ensure_keep(c.lvl_first, nc * 4, used * 4, c.stream));
    SVO_TRY(ensure(c.lvl_start, nc * 8));
    cap = nc;
    return hipSuccess;
  };
  auto grow = [&](uint64_t need) -> hipError_t {
    if (need <= cap) return hipSuccess;
    return alloc_levels(std::max(need, cap * 2));
  };
  SVO_TRY(alloc_levels(cap));
  ull* ctr = c.counters.as<ull>();
  ull hn = 0;
  for (int l = D - 1; l >= 0; l--) {
    const uint64_t nch = n[l + 1];
    SVO_TRY(grow(used + nch));
    off[l] = used;
    LevelView ch;
    ch.code = (l + 1 == D) ? c.codes.as<uint64_t>() : c.lvl_code.as<uint64_t>() + off[l + 1];
    ch.size = (l + 1 == D) ? nullptr : c.lvl_size.as<uint64_t>() + off[l + 1];
    ch.n = nch;
    ParentStore st{ch,
                   c.lvl_code.as<uint64_t>() + used,
                   c.lvl_size.as<uint64_t>() + used,
                   c.lvl_mask.as<uint8_t>() + used,
                   c.lvl_first.as<uint32_t>() + used,
                   ctr + kCtrLevelN};
    SVO_TRY(ensure(c.scan_scratch, scan_scratch_elems(nch) * 4 + 256));
    SVO_TRY((scan_exclusive<uint32_t>(HeadLoad{ch.code}, st, nch,
                                      c.scan_scratch.as<uint32_t>(), c.stream)));
    SVO_TRY(hipMemcpyAsync(&hn, ctr + kCtrLevelN, 8, hipMemcpyDeviceToHost, c.stream));
    SVO_TRY(hipStreamSynchronize(c.stream));
    n[l] = hn;
    used += hn;
  }
  // root level must hold exactly the root
  if (n[0] != 1) {
    set_error("internal: root level does not hold one node");
    return hipErrorUnknown;
  }
  uint64_t size_root = 0;
  SVO_TRY(hipMemcpyAsync(&size_root, c.lvl_size.as<uint64_t>() + off[0], 8,
                         hipMemcpyDeviceToHost, c.stream));
  SVO_TRY(hipStreamSynchronize(c.stream));
  c.n_nodes = size_root;
  SVO_TRY(ensure(c.nodes, size_root * 24));

  // ---- top-down
  uint64_t* start = c.lvl_start.as<uint64_t>();
  hipLaunchKernelGGL(k_emit_root, dim3(1), dim3(1), 0, c.stream, c.lvl_size.as<uint64_t>() + off[0],
                     c.lvl_mask.as<uint8_t>() + off[0], start + off[0], c.nodes.as<uint64_t>());
  SVO_TRY(hipGetLastError());
  for (int l = 0; l < D; l++) {
    const bool leaf = (l + 1 == D);
    const uint64_t np = n[l];
    ktimer_begin(c, "emit_children");
    hipLaunchKernelGGL(k_emit_children, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c.stream,
                       np, start + off[l], c.lvl_size.as<uint64_t>() + off[l],
                       c.lvl_mask.as<uint8_t>() + off[l], c.lvl_first.as<uint32_t>() + off[l],
                       leaf ? nullptr : c.lvl_size.as<uint64_t>() + off[l + 1],
                       leaf ? nullptr : c.lvl_mask.as<uint8_t>() + off[l + 1],
                       leaf ? nullptr : start + off[l + 1], leaf ? 1 : 0, c.payload, 1ull,
                       c.nodes.as<uint64_t>());
    ktimer_end(c);
    SVO_TRY(hipGetLastError());
  }
  return hipStreamSynchronize(c.stream);
}

}  // namespace svo
