// Streamed (out-of-core) SVO build: the octree assembled partition by partition, the way
// the reference's OctreeBuilder streams its records to disk while main.cpp feeds it one
// Morton-range partition at a time (main.cpp:329-387, OctreeBuilder.cpp:4-168).
//
// Partition i is an aligned Morton cube, so its voxels form the subtree of one node Q_i
// at depth log8(P). The streaming builder writes every node record of that subtree except
// Q_i itself while partition i is fed, before anything of partition i+1; Q_i is written
// when its group of 8 sibling partitions completes (after partition 8j+7), the group of
// their parents when 64 partitions complete, and so on; the root comes last
// (OctreeBuilder.cpp:34-55, 58-79, 112-128). With P = 8:
//   nodes = [subtree(Q_0) - Q_0] ... [subtree(Q_7) - Q_7] [non-null Q_0..Q_7] root
// Each partition is built alone by the level passes (svo_build.hip) over its own voxels
// at the global gridsize. That tree is Q_i's subtree plus the chain of D - Dp single-child
// ancestors up to the root, and in its post-order output Q_i sits at size(Q_i) - 1 with
// the chain after it. Records [0, size(Q_i) - 1) are the partition's output once their
// children_base moves by the node records emitted before, and their data index by the
// data records emitted before (the 0 of internal nodes and the shared white record 1 of
// geometry-only leaves stay); the build writes them so moved, straight after the records
// emitted before (or to the start of its own array, for the sink), so no record is
// copied again.
//
// Data records follow the same order: payload voxel records in voxel order; with -levels
// every record in post-order, so partition i contributes its subtree's records (Q_i's
// average last) and the chain's records are dropped. The partition roots and the levels
// above them (P = 8^k nodes at most per level) are grouped on the host, as the
// reference's buffers group eight siblings (OctreeBuilder.cpp:58-128), each group's
// records going out when it completes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "svo_internal.h"

namespace svo {

#define SVO_TRY(x)                   \
  do {                               \
    hipError_t _e = (x);             \
    if (_e != hipSuccess) return _e; \
  } while (0)

constexpr uint64_t kNoChildren = ~0ull;  // offsets[8] all -1: a leaf

static int ilog2(uint64_t v) {
  int d = 0;
  while ((1ull << d) < v) d++;
  return d;
}

static hipError_t sink_error(Context& c) {
  c.st.sink_failed = true;
  set_error("the output sink stopped the build");
  return hipErrorUnknown;
}

// Device bytes to the sink: chunks of whole records copied into two pinned buffers, the
// copy of one chunk running while the sink consumes the other.
static hipError_t drain_to_sink(Context& c, int kind, const void* dsrc, uint64_t bytes,
                                uint64_t rec) {
  const size_t cap = 64u << 20;
  if (c.drain_cap < cap) {
    for (int k = 0; k < 2; k++) {
      if (c.drain[k]) (void)hipHostFree(c.drain[k]);
      c.drain[k] = nullptr;
    }
    c.drain_cap = 0;
    for (int k = 0; k < 2; k++) {
      SVO_TRY(hipHostMalloc(&c.drain[k], cap, hipHostMallocDefault));
      if (!c.drain_ev[k]) SVO_TRY(hipEventCreateWithFlags(&c.drain_ev[k], hipEventDisableTiming));
    }
    c.drain_cap = cap;
  }
  const uint64_t chunk = (c.drain_cap / rec) * rec;
  uint64_t pend[2] = {0, 0};
  const char* src = (const char*)dsrc;
  uint64_t off = 0;
  for (uint64_t k = 0; off < bytes; k++) {
    const int b = (int)(k & 1);
    if (pend[b]) {  // chunk k - 2: copied by now or soon; the sink takes it before reuse
      SVO_TRY(event_sync(c.drain_ev[b]));
      if (c.st.sink(c.st.user, kind, c.drain[b], pend[b]) != 0) return sink_error(c);
      pend[b] = 0;
    }
    const uint64_t n = std::min<uint64_t>(chunk, bytes - off);
    SVO_TRY(hipMemcpyAsync(c.drain[b], src + off, n, hipMemcpyDeviceToHost, c.stream));
    SVO_TRY(hipEventRecord(c.drain_ev[b], c.stream));
    pend[b] = n;
    off += n;
  }
  // the two chunks still held, older first
  const uint64_t nk = (bytes + chunk - 1) / chunk;
  for (uint64_t k = nk >= 2 ? nk - 2 : 0; k < nk; k++) {
    const int b = (int)(k & 1);
    if (!pend[b]) continue;
    SVO_TRY(event_sync(c.drain_ev[b]));
    if (c.st.sink(c.st.user, kind, c.drain[b], pend[b]) != 0) return sink_error(c);
    pend[b] = 0;
  }
  return hipSuccess;
}

// Host records (the leading data records, the levels above the partitions) to the
// output: straight to the sink, or appended to the device arrays.
static hipError_t emit_host(Context& c, int kind, const void* src, uint64_t bytes) {
  if (!bytes) return hipSuccess;
  StreamState& s = c.st;
  if (s.sink) {
    if (s.sink(s.user, kind, src, bytes) != 0) return sink_error(c);
    return hipSuccess;
  }
  DevBuf& b = kind == 0 ? s.out_nodes : s.out_data;
  const uint64_t used = kind == 0 ? s.nodes_out * 24 : s.data_out * 32;
  SVO_TRY(ensure_keep(b, used + bytes, used, c.stream));
  return upload_async(c, (char*)b.p + used, src, bytes);
}

static void reset_voxel_stream(Context& c) {
  c.n_vox = 0;
  c.n_bricks = 0;
  c.bricks_ok = true;
  c.codes_lazy = false;
}

hipError_t stream_begin(Context& c, uint64_t gridsize, uint64_t n_parts, int levels,
                        SinkFn sink, void* user) {
  StreamState& s = c.st;
  s.active = true;
  s.gridsize = gridsize;
  s.n_parts = n_parts;
  s.levels = levels ? 1 : 0;
  s.payload = c.payload;
  s.depth = ilog2(gridsize);
  s.part_depth = s.depth - ilog2(n_parts) / 3;
  s.sink = sink;
  s.user = user;
  s.next_part = 0;
  s.nodes_out = s.data_out = 0;
  s.tbuf.assign((size_t)(ilog2(n_parts) / 3) + 1, std::vector<TopNode>());
  s.any_voxels = false;
  s.hazard = 0;
  s.sink_failed = false;
  reset_voxel_stream(c);
  c.n_nodes = c.n_data = 0;
  if (!sink) {
    // the larger of the two node (data) arrays, usually the previous run's output,
    // collects this run's records (the partitions' builds write their node records into
    // it; their data records go through the other): no growth copies when a mesh is built
    // again
    if (c.nodes.cap > s.out_nodes.cap) std::swap(c.nodes, s.out_nodes);
    if (c.records.cap > s.out_data.cap) std::swap(c.records, s.out_data);
  }
  // OctreeBuilder.cpp:25-29: the null record, and for geometry-only builds the white
  // record every leaf refers to
  float hdr[16];
  memset(hdr, 0, sizeof(hdr));
  hdr[10] = hdr[11] = hdr[12] = 1.0f;
  const uint64_t nh = s.payload ? 1 : 2;
  SVO_TRY(emit_host(c, 1, hdr, nh * 32));
  s.data_out = nh;
  return hipSuccess;
}

static inline float canon_nan_host(float f) {
  if (f != f) {
    const uint32_t u = 0xFFC00000u;
    memcpy(&f, &u, 4);
  }
  return f;
}

// A completed group of 8 siblings (OctreeBuilder.cpp:58-102, 112-128): a null parent when
// all are null; otherwise the non-null ones are written consecutively, the parent points
// at the first, and with -levels its average is written as a data record.
static hipError_t top_group(Context& c, const std::vector<TopNode>& ch, TopNode& p) {
  StreamState& s = c.st;
  bool any = false;
  for (const TopNode& t : ch) any = any || !t.null;
  p = TopNode();
  if (!any) return hipSuccess;
  p.null = false;
  uint64_t recs[8 * 3];
  int nrec = 0;
  uint64_t cb = 0, offs = 0;
  for (int k = 0; k < 8; k++) {
    uint64_t byte = 0xFF;
    if (!ch[k].null) {
      const uint64_t pos = s.nodes_out + (uint64_t)nrec;
      memcpy(recs + 3 * nrec, ch[k].rec, 24);
      if (nrec == 0) cb = pos;
      byte = pos - cb;
      nrec++;
    }
    offs |= byte << (8 * k);
  }
  SVO_TRY(emit_host(c, 0, recs, (uint64_t)nrec * 24));
  s.nodes_out += (uint64_t)nrec;
  p.rec[0] = 0;
  p.rec[1] = cb;
  p.rec[2] = offs;
  if (s.levels) {  // OctreeBuilder.cpp:82-99, the float order of average_children
    float a[6] = {0, 0, 0, 0, 0, 0};
    float notnull = 0.0f;
    for (int k = 0; k < 8; k++) {
      if (!ch[k].null) notnull += 1.0f;
      for (int q = 0; q < 6; q++) a[q] = a[q] + ch[k].cache[q];
    }
    float avg[6];
    avg[0] = a[0] / notnull;
    avg[1] = a[1] / notnull;
    avg[2] = a[2] / notnull;
    const float t0 = a[3] / notnull, t1 = a[4] / notnull, t2 = a[5] / notnull;
    const float inv = 1.0f / sqrtf((t0 * t0 + t1 * t1) + t2 * t2);
    avg[3] = t0 * inv;
    avg[4] = t1 * inv;
    avg[5] = t2 * inv;
    float rec[8];
    memset(rec, 0, 8);  // morton 0
    for (int q = 0; q < 6; q++) rec[2 + q] = canon_nan_host(avg[q]);
    SVO_TRY(emit_host(c, 1, rec, 32));
    p.rec[0] = s.data_out++;
    memcpy(p.cache, avg, sizeof(avg));
  }
  return hipSuccess;
}

// The next partition root (or null) at the partition level; full groups of 8 go up.
static hipError_t top_push(Context& c, const TopNode& t) {
  StreamState& s = c.st;
  int L = (int)s.tbuf.size() - 1;
  s.tbuf[L].push_back(t);
  while (L > 0 && s.tbuf[L].size() == 8) {
    TopNode p;
    SVO_TRY(top_group(c, s.tbuf[L], p));
    s.tbuf[L].clear();
    L--;
    s.tbuf[L].push_back(p);
  }
  return hipSuccess;
}

// Partitions [next_part, end) hold no voxels: null roots.
static hipError_t top_skip_to(Context& c, uint64_t end) {
  StreamState& s = c.st;
  for (; s.next_part < end; s.next_part++) SVO_TRY(top_push(c, TopNode()));
  return hipSuccess;
}

hipError_t stream_partition(Context& c, uint64_t part) {
  StreamState& s = c.st;
  SVO_TRY(top_skip_to(c, part));
  s.next_part = part + 1;
  if (c.n_vox == 0) {  // no voxels: a null subtree, nothing written
    reset_voxel_stream(c);
    return top_push(c, TopNode());
  }
  s.any_voxels = true;
  Job& J = c.jobs[0];
  J.gridsize = s.gridsize;
  J.voff = 0;
  J.nvox = c.n_vox;
  // the build writes the partition's records already placed: appended to the collected
  // node records, or from 0 of the build's own array when they go to the sink
  const uint64_t hdr = s.payload ? 1 : 2;
  const BuildOut out{s.sink ? &c.nodes : &s.out_nodes, s.sink ? 0 : s.nodes_out, s.nodes_out,
                     s.data_out - hdr};
  SVO_TRY(build_jobs(c, 1, false, s.levels, &out));
  if (c.err_count) {
    set_error("internal: " + std::to_string(c.err_count) +
              " consistency failures in partition " + std::to_string(part) +
              " (scan look-back time-outs, brick pool slots not found or payload voxels "
              "without a covering triangle)");
    return hipErrorUnknown;
  }
  s.hazard |= J.hazard;
  const uint64_t chain = (uint64_t)(s.depth - s.part_depth);
  if (c.n_nodes < chain + 1) {
    set_error("internal: a partition's tree is smaller than its ancestor chain");
    return hipErrorUnknown;
  }
  const uint64_t q = c.n_nodes - 1 - chain;  // Q's record; [0, q) is the partition's output
  const uint64_t* placed = out.nodes->as<uint64_t>() + 3 * out.at;
  TopNode pr;
  pr.null = false;
  {
    const ReadReq rq = {pr.rec, placed + 3 * q, 24};
    SVO_TRY(read_back(c, &rq, 1));
  }
  // -levels: Q's average (an internal node's record; a payload leaf's own voxel record;
  // geometry-only leaves carry zero caches), read at its index within the partition
  if (s.levels && (pr.rec[2] != kNoChildren || s.payload)) {
    const uint64_t di = pr.rec[0] >= hdr ? pr.rec[0] - out.data_add : pr.rec[0];
    float rec[8];
    const ReadReq rq = {rec, (const char*)c.data_out().p + di * 32, 32};
    SVO_TRY(read_back(c, &rq, 1));
    memcpy(pr.cache, rec + 2, sizeof(pr.cache));
  }
  // records [q, n_nodes) (Q and its ancestor chain) lie past the partition's output: the
  // next records emitted overwrite them
  if (q && s.sink) SVO_TRY(drain_to_sink(c, 0, placed, q * 24, 24));
  uint64_t d0 = 0, dn = 0;
  if (s.levels) {
    d0 = hdr;
    dn = c.n_data - hdr - chain;
  } else if (s.payload) {
    d0 = 1;
    dn = c.n_vox;
  }
  if (dn) {
    const char* src = (const char*)c.data_out().p + d0 * 32;
    if (s.sink) {
      SVO_TRY(drain_to_sink(c, 1, src, dn * 32, 32));
    } else {
      SVO_TRY(ensure_keep(s.out_data, (s.data_out + dn) * 32, s.data_out * 32, c.stream));
      SVO_TRY(hipMemcpyAsync((char*)s.out_data.p + s.data_out * 32, src, dn * 32,
                             hipMemcpyDeviceToDevice, c.stream));
    }
  }
  s.nodes_out += q;
  s.data_out += dn;
  reset_voxel_stream(c);
  return top_push(c, pr);
}

hipError_t stream_end(Context& c, uint64_t* n_nodes, uint64_t* n_data, int* hazard) {
  StreamState& s = c.st;
  SVO_TRY(top_skip_to(c, s.n_parts));  // the remaining partitions are empty
  // the root, last (OctreeBuilder.cpp:42); an empty tree's root is a null record
  const TopNode& root = s.tbuf[0].at(0);
  SVO_TRY(emit_host(c, 0, root.rec, 24));
  s.nodes_out++;
  const uint64_t nn = s.nodes_out, nd = s.data_out;
  // an empty 1^3 grid: the reference reads a root buffer that was never filled
  if (s.gridsize == 1 && !s.any_voxels) s.hazard = 1;
  s.active = false;
  if (s.sink) {
    SVO_TRY(stream_sync(c.stream));
    c.n_nodes = c.n_data = 0;  // nothing of the octree is held on the device
  } else {
    std::swap(c.nodes, s.out_nodes);
    std::swap(c.records, s.out_data);
    c.data_in_records2 = false;
    c.n_nodes = nn;
    c.n_data = nd;
  }
  c.hazard = s.hazard;
  Job& J = c.jobs[0];
  J.n_nodes = nn;
  J.n_data = nd;
  J.nbase = J.dbase = 0;
  J.hazard = s.hazard;
  if (n_nodes) *n_nodes = nn;
  if (n_data) *n_data = nd;
  if (hazard) *hazard = s.hazard;
  return hipSuccess;
}

void stream_release(Context& c) {
  release(c.st.out_nodes);
  release(c.st.out_data);
  for (int k = 0; k < 2; k++) {
    if (c.drain[k]) (void)hipHostFree(c.drain[k]);
    if (c.drain_ev[k]) (void)hipEventDestroy(c.drain_ev[k]);
    c.drain[k] = nullptr;
    c.drain_ev[k] = nullptr;
  }
  c.drain_cap = 0;
}

}  // namespace svo
