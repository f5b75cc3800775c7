// Library-internal state and stage drivers behind the C ABI (include/svo_amd.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace svo {

// Kernel launches and host synchronizations issued by this thread (a context is used by
// one thread at a time, so the difference across a call is that call's count). Launches
// recorded into a graph count as graph nodes; t_graph_launches counts graph launches,
// t_graph_replays those of a graph kept from an earlier call.
inline thread_local uint64_t t_launches = 0;
inline thread_local uint64_t t_syncs = 0;
inline thread_local uint64_t t_graph_nodes = 0;
inline thread_local uint64_t t_graph_launches = 0;
inline thread_local uint64_t t_graph_replays = 0;
inline thread_local uint64_t t_graph_kernels = 0;  // kernels inside the launched graphs
inline thread_local uint64_t t_dry_launches = 0;   // launches recorded by dry passes
inline hipError_t stream_sync(hipStream_t s) {
  t_syncs++;
  return hipStreamSynchronize(s);
}
inline hipError_t event_sync(hipEvent_t e) {
  t_syncs++;
  return hipEventSynchronize(e);
}

// Launch modes of this thread: kLaunchRun launches, kLaunchCapture launches into a stream
// capture (graph nodes), kLaunchDry only records. In the last two every launch appends its
// kernel, grid, block, shared memory, stream and argument bytes (as passed to the kernel)
// to t_launch_log: two segments of launches with equal logs make identical graphs.
enum LaunchMode : int { kLaunchRun = 0, kLaunchCapture = 1, kLaunchDry = 2 };
inline thread_local int t_launch_mode = kLaunchRun;
inline thread_local std::vector<uint8_t>* t_launch_log = nullptr;

template <typename T>
inline void log_bytes(std::vector<uint8_t>& L, const T& v) {
  const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
  L.insert(L.end(), p, p + sizeof(T));
}

template <typename... KArgs, typename... Args, size_t... I>
inline hipError_t launch_impl(void (*k)(KArgs...), dim3 g, dim3 b, uint32_t sh, hipStream_t s,
                              std::index_sequence<I...>, Args&&... args) {
  // the arguments converted to the kernel's parameter types, as the kernel receives them
  std::tuple<KArgs...> t{static_cast<KArgs>(std::forward<Args>(args))...};
  if (t_launch_log) {
    std::vector<uint8_t>& L = *t_launch_log;
    log_bytes(L, reinterpret_cast<const void*>(k));
    log_bytes(L, g);
    log_bytes(L, b);
    log_bytes(L, sh);
    log_bytes(L, s);
    (log_bytes(L, std::get<I>(t)), ...);
  }
  if (t_launch_mode == kLaunchDry) {
    t_dry_launches++;
    return hipSuccess;
  }
  if (t_launch_mode == kLaunchCapture)
    t_graph_nodes++;
  else
    t_launches++;
  void* ptrs[sizeof...(KArgs) + 1] = {reinterpret_cast<void*>(&std::get<I>(t))..., nullptr};
  return hipLaunchKernel(reinterpret_cast<const void*>(k), g, b, ptrs, sh, s);
}

template <typename... KArgs, typename... Args>
inline hipError_t launch_kernel(void (*k)(KArgs...), dim3 g, dim3 b, uint32_t sh, hipStream_t s,
                                Args&&... args) {
  static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
  return launch_impl(k, g, b, sh, s, std::index_sequence_for<KArgs...>{},
                     std::forward<Args>(args)...);
}

// every kernel launch of the library goes through this (errors reach hipGetLastError)
#define SVO_LAUNCH(...) (void)::svo::launch_kernel(__VA_ARGS__)

// Sparse occupancy grid: the partition's Morton range is cut into aligned bricks of
// 2^kBrickBits voxels (16^3). A dense table maps brick -> pool slot (0 = empty). A slot
// holds the brick's voxels as bytes (marked by plain stores) and, after compaction, as
// 64 bit words, both in Morton order.
constexpr int kBrickBits = 12;
constexpr int kBrickVox = 1 << kBrickBits;
constexpr int kBrickWords = 1 << (kBrickBits - 6);
// Per-brick counts carry a non-empty flag at bit kNonEmptyShift above the voxel count.
constexpr int kNonEmptyShift = 40;
constexpr uint64_t kVoxMask = (1ull << kNonEmptyShift) - 1;

// Device counters block (one 256-byte line of u64s).
enum Counter : int {
  kCtrPoolCount = 0,  // bricks allocated from the pool
  kCtrOverflow = 1,   // pool exhausted during a pass
  kCtrLargeCount = 2, // triangles deferred to the column-parallel tier
  kCtrNumOcc = 3,     // occupied bricks after compaction
  kCtrNumVox = 4,     // voxels of the current pass
  kCtrLevelN = 5,     // node count of the level being built
  kCtrCols = 6,       // columns of the large tier
  kCtrError = 15,     // sticky: internal consistency failures (cleared per job)
  kCtrPartTotal = 16, // partition list length
  kNumCounters = 32
};
constexpr int kCtrPassClear = 15;  // counters 0..14 are cleared at the start of a pass

// Voxel codes of a batch carry their job index in bits kJobShift and up, so that one
// sorted stream holds the voxels of many meshes (job-major, Morton order within a job).
// A job's gridsize is then at most 2^16 (codes below 2^48); a single job uses tag 0
// and any gridsize.
constexpr int kJobShift = 48;
constexpr uint64_t kCodeMask = (1ull << kJobShift) - 1;
constexpr uint32_t kMaxJobs = 1u << 16;

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

// Voxelizer tuning: small/large split threshold (columns per triangle) and block size.
struct VoxTuning {
  int small_cols = 1024;    // columns above which a triangle goes to the large tier
  int direct_cols = 16;     // columns up to which a triangle's own thread walks them
  int tile_blocks = 1024;   // persistent grid of the small tier
  int large_blocks = 2048;  // grid of the large tier (grid-stride)
};

struct StageTimes {
  float partition_ms = 0, voxelize_ms = 0, build_ms = 0;
};

// Per-kernel device time, measured with HIP events on the context's stream when kernel
// timing is switched on (the counterpart of the reference's stage timers, globals.h:8-29).
struct KernelStat {
  uint64_t launches = 0;
  double ms = 0;
};
struct PendingTimer {
  std::string name;
  hipEvent_t a, b;
};

// One mesh to voxelize and build. The plain flow (svo_mesh_load, svo_partition,
// svo_voxelize, svo_build) uses job 0; a batch uses one job per mesh.
struct Job {
  const float* tris = nullptr;  // device records (9 or 21 floats per triangle)
  uint64_t n_tris = 0;
  int payload = 0;
  int color_mode = 0;
  uint64_t gridsize = 0;
  float unit_vox = 0.0f;  // voxelizer unitlength (main.cpp:311)
  // partitions: lists in Context::part_lists at part_off[i], part_cnt[i] triangles
  uint64_t n_parts = 0;
  std::vector<uint64_t> part_off, part_cnt;
  // results
  uint64_t voff = 0, nvox = 0;    // range of the voxel stream (Context::codes)
  uint64_t dbase = 0, n_data = 0; // range of the data records (Context::records)
  uint64_t nbase = 0, n_nodes = 0;// range of the node records (Context::nodes)
  int hazard = 0;
};

// Device view of a job for the payload records and the build.
struct JobDev {
  const float* tris;
  float unit_div;
  int color_mode;
  int payload;
  int depth;  // log2 gridsize
  int levels; // -levels: internal nodes carry averaged data records
  uint64_t gridsize;
  uint64_t voff, nvox, dbase;
  uint64_t n_nodes, nbase, hazard, n_data;  // written by the build
  // placement of the records in a larger node stream (a streamed partition; 0 otherwise):
  // children_base of nodes with children moves by cb_add, data indices past the leading
  // records (null, and white for geometry-only builds) by data_add
  uint64_t cb_add, data_add;
};

// Host description of one voxelize unit (one partition of one job).
struct VoxUnit {
  uint32_t job;
  const float* tris;
  const uint32_t* list;  // nullptr = all triangles of the job
  uint64_t n;
  int stride;
  uint64_t mstart, mend;
  float unit;
  uint64_t gridsize;
};

// Output sink of the streamed build: consecutive chunks of the .octreenodes (kind 0) and
// .octreedata (kind 1) byte streams, in host memory valid during the call (the C ABI's
// svo_sink_fn). A non-zero return stops the build.
typedef int (*SinkFn)(void* user, int kind, const void* bytes, uint64_t size);

// A node of the levels at and above the partition roots: its 24-byte record
// (children_base and data index already global) and, with -levels, its averaged colour
// and normal; null = no voxels below.
struct TopNode {
  bool null = true;
  uint64_t rec[3] = {0, 0, ~0ull};
  float cache[6] = {0, 0, 0, 0, 0, 0};
};

// Streamed (out-of-core) build, OctreeBuilder.cpp:4-168 driven per partition as
// main.cpp:329-387 does: each partition's subtree is built alone and its records leave
// (to the sink, or appended to out_nodes / out_data in HBM) before the next partition;
// the partition roots and the levels above them are grouped by eight as their groups
// complete, and the root comes last.
struct StreamState {
  bool active = false;
  uint64_t gridsize = 0, n_parts = 0;
  int levels = 0, payload = 0;
  int depth = 0, part_depth = 0;  // log2 gridsize, log2 of a partition's edge
  SinkFn sink = nullptr;
  void* user = nullptr;
  uint64_t next_part = 0;               // partitions below this one are done
  uint64_t nodes_out = 0, data_out = 0; // records emitted so far
  std::vector<std::vector<TopNode>> tbuf;  // per level 0 (root) .. log8(P) (partition roots)
  bool any_voxels = false;
  int hazard = 0;
  bool sink_failed = false;             // the sink returned non-zero
  DevBuf out_nodes, out_data;           // the records, when there is no sink
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  VoxTuning tune;

  // mesh of the plain flow (device-resident records, 9 or 21 floats each)
  DevBuf tris;
  uint64_t n_tris = 0;
  int payload = 0;

  // jobs: job 0 is the plain flow's mesh; after a batch, the batch's meshes
  std::vector<Job> jobs;
  bool batch_state = false;  // the outputs are a batch's
  DevBuf jobdev;  // JobDev[jobs.size()]
  DevBuf batch_tris;  // records of a batch's meshes loaded from host memory

  // partitioning (lists of all jobs' partitions, concatenated)
  DevBuf hitbox, part_lists;
  uint64_t part_used = 0;

  // voxelizer scratch (kept zeroed between passes: table, grid8, pool8)
  DevBuf unitdev;  // VoxParams[n] | ubase[n+1] | tile_off[n+1] | uvoff[n+1]
  DevBuf payload_map;  // a pass's payload units: voxel offsets, first voxels, unit indices
  std::vector<uint8_t> unit_img;  // what unitdev holds before uvoff (empty: unknown)
  DevBuf table, pool, pool8;  // pool: bit words per slot; pool8: bytes per slot
  DevBuf grid8;               // dense byte grids of a pass's units (when they fit)
  uint64_t table_entries = 0;
  uint64_t grid_zeroed = 0;     // bytes of grid8 known to be zero
  uint64_t pool_cap = 0;        // slots of pool8 (and of pool) on the touch path
  uint64_t dense_limit = 0;     // largest dense grid in bytes (set from free HBM)
  DevBuf counters;
  DevBuf large_unit, large_tri, large_cols, large_off;
  DevBuf large_walk;  // walk constants of large triangles (44 words each, capped)
  DevBuf occ_list, bcount, wprefix, winner;
  DevBuf scan_scratch;
  DevBuf scan_status;  // status words of the single-pass scans of a voxelize pass

  // brick stream: every occupied 16^3 brick of the voxel stream, in stream order: its 64
  // occupancy words, its first voxel's code (tagged) and its first voxel's stream index.
  // The build derives the four lowest tree levels from it.
  DevBuf bwords, bcode, bvoff;
  DevBuf bshape;  // per brick: subtree size and child mask of its node (brick_shape_word)
  uint64_t n_bricks = 0;
  bool bricks_ok = true;  // every voxel of the stream came with its brick, bricks unique
  // Geometry-only passes leave the codes of their voxels unwritten (the build reads the
  // brick stream); ensure_codes expands the brick stream into them when a reader needs them.
  bool codes_lazy = false;

  // voxel stream across partitions and jobs (sorted by construction)
  DevBuf codes;    // u64 Morton codes (job tag in the high bits for batches)
  DevBuf records;  // 32 B data records of every job (each job: null record first)
  uint64_t n_vox = 0;
  uint64_t nfilled = 0;

  // build output
  DevBuf nodes;
  uint64_t n_nodes = 0, n_data = 0;
  int hazard = 0;
  unsigned long long err_count = 0;  // the sticky failure counter, read by the build
  DevBuf lvl_code, lvl_size, lvl_mask, lvl_first, lvl_start, lvl_aux, lvl_leafp, lvl_pidx;
  DevBuf lvl_sib;  // per arena node: sum of (size - 1) over its preceding siblings
  DevBuf up_tiles; // upper levels at once: per tile and height counts, job root nodes
  std::vector<JobDev> jobdev_img;  // what jobdev holds on the device (empty: unknown)
  DevBuf lvl_isize, lvl_cache, lvl_pstart;  // -levels: internal counts, averages, data starts
  DevBuf records2;                          // -levels: data records in post-order
  bool data_in_records2 = false;            // the built data records are in records2
  const DevBuf& data_out() const { return data_in_records2 ? records2 : records; }

  StreamState st;
  // helper context (an svo_ctx) of a batch's -levels jobs, created on first use
  void* aux = nullptr;
  // streamed output: two pinned chunk buffers for the device-to-host copies that feed the
  // sink, and their copy events
  void* drain[2] = {nullptr, nullptr};
  size_t drain_cap = 0;
  hipEvent_t drain_ev[2] = {nullptr, nullptr};

  // streaming mesh ingest: two pinned chunk buffers and their copy events
  void* ingest[2] = {nullptr, nullptr};
  size_t ingest_cap = 0;
  hipEvent_t ingest_ev[2] = {nullptr, nullptr};

  // pinned host staging for small uploads (bump allocated; wraps after a stream sync)
  void* stage = nullptr;
  size_t stage_cap = 0, stage_used = 0;
  // pinned host landing area for device-to-host reads (counters, layouts); its first
  // 256 bytes hold the completion flag the gather kernel raises
  void* hread = nullptr;
  size_t hread_cap = 0;
  uint32_t hread_seq = 0;

  StageTimes times;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  hipEvent_t build_done_ev = nullptr;  // when set, the build records it before its read-back

  bool ktiming = false;
  std::vector<PendingTimer> kpending;
  std::vector<hipEvent_t> evpool;
  std::map<std::string, KernelStat> kstats;

  // graph segments (run_segment): while 'graphs' is set (a batch with graph_policy on),
  // each kernel sequence between two host synchronizations is captured into a graph; the
  // graph is kept, with the launch log it was captured from, under its segment index
  // (gseg counts the segments of the current call), and launched again whenever a later
  // call makes the same launches at that index
  bool graphs = false;
  int graph_policy = 1;
  uint32_t gseg = 0;
  std::vector<std::vector<uint8_t>> glog;
  std::vector<hipGraphExec_t> gexec;
};

// Graph segments on for the lifetime of a scope (one batch call), numbered from 0.
struct GraphScope {
  Context& c;
  GraphScope(Context& ctx, bool on) : c(ctx) {
    c.graphs = on;
    c.gseg = 0;
  }
  ~GraphScope() { c.graphs = false; }
};

// Run body (a sequence of SVO_LAUNCHes and nothing else that reaches the device; host
// effects only where repeating them is harmless) as one graph segment of c: a dry pass
// records its launch log; a kept graph with an equal log is launched in its place,
// otherwise body runs again under stream capture and its graph replaces the kept one.
// Without c.graphs (or while kernels are timed, or on the null stream) body just runs.
template <typename F>
hipError_t run_segment(Context& c, F&& body) {
  if (!c.graphs || c.ktiming || !c.stream || t_launch_mode != kLaunchRun) return body();
  const uint32_t idx = c.gseg++;
  if (c.glog.size() <= idx) {
    c.glog.resize(idx + 1);
    c.gexec.resize(idx + 1, nullptr);
  }
  std::vector<uint8_t> log;
  log.reserve(c.glog[idx].size());
  const uint64_t d0 = t_dry_launches;
  t_launch_mode = kLaunchDry;
  t_launch_log = &log;
  hipError_t e = body();
  t_launch_mode = kLaunchRun;
  t_launch_log = nullptr;
  if (e != hipSuccess) return e;
  const uint64_t n_kernels = t_dry_launches - d0;
  if (n_kernels == 0) return hipSuccess;  // nothing to launch
  if (c.gexec[idx] && log == c.glog[idx]) {
    t_graph_launches++;
    t_graph_replays++;
    t_graph_kernels += n_kernels;
    return hipGraphLaunch(c.gexec[idx], c.stream);
  }
  log.clear();
  e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) return e;
  t_launch_mode = kLaunchCapture;
  t_launch_log = &log;
  const hipError_t eb = body();
  t_launch_mode = kLaunchRun;
  t_launch_log = nullptr;
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(c.stream, &g);
  if (eb != hipSuccess || ee != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return eb != hipSuccess ? eb : ee;
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return e;
  // the kept graph was last launched by an earlier call, which waited for its stream
  if (c.gexec[idx]) (void)hipGraphExecDestroy(c.gexec[idx]);
  c.gexec[idx] = x;
  c.glog[idx].swap(log);
  t_graph_launches++;
  t_graph_kernels += n_kernels;
  return hipGraphLaunch(x, c.stream);
}

// Kernel timing helpers (no-ops unless c.ktiming): begin/end bracket one launch on
// c.stream; flush folds finished pairs into c.kstats after a stream synchronization.
void ktimer_begin(Context& c, const char* name);
void ktimer_end(Context& c);
void ktimer_flush(Context& c);

// Allocation helpers: 'ensure' does not preserve contents; 'ensure_keep' preserves the
// first 'used' bytes.
hipError_t ensure(DevBuf& b, size_t bytes);
hipError_t ensure_keep(DevBuf& b, size_t bytes, size_t used, hipStream_t s);
void release(DevBuf& b);

// Copy 'bytes' from host memory to device memory 'dst' asynchronously on c.stream
// through the pinned staging buffer (the source may be reused at once).
hipError_t upload_async(Context& c, void* dst, const void* src, size_t bytes);

// Read device regions into host memory with one stream synchronization, through the
// pinned landing area (a direct copy, not the staged path of pageable memory).
struct ReadReq {
  void* dst;
  const void* src;
  size_t bytes;
};
hipError_t read_back(Context& c, const ReadReq* reqs, int n);

// Stage drivers (return hipSuccess or an error; message in last error).
// Partition job j's mesh into n_parts Morton-range partitions; lists appended to
// c.part_lists (reset c.part_used first to start over).
hipError_t partition_job(Context& c, uint32_t j, float unitlength, uint64_t gridsize,
                         uint64_t n_parts);
// out[k] = the record (stride floats) of triangle list[k], or of triangle k when list is
// null, for k < n, on c.stream.
hipError_t gather_records(Context& c, const float* tris, int stride, const uint32_t* list,
                          uint64_t n, float* out);
// Voxelize the units in the same launches and append their voxels to the stream.
// batch = false: every unit belongs to job 0, whose voxels accumulate over calls;
// batch = true: the units hold the whole of their jobs, whose ranges are laid out here.
hipError_t voxelize_units(Context& c, const std::vector<VoxUnit>& units, bool batch,
                          uint64_t* nfilled);
// Write the codes of every voxel that only the brick stream holds (no-op when none).
hipError_t ensure_codes(Context& c);
// Where a single build's node records go instead of Context::nodes from 0: the array
// 'nodes' from record 'at' on (its first 'at' records are kept when it grows), with the
// record fields moved by cb_add and data_add (JobDev).
struct BuildOut {
  DevBuf* nodes;
  uint64_t at;
  uint64_t cb_add, data_add;
};
// Build the octrees of jobs [0, n_jobs) from the voxel stream; 'tagged' = the stream
// carries job tags (a batch). levels (-levels, OctreeBuilder.cpp:82-99) is supported for
// a single job, and so is 'out'.
hipError_t build_jobs(Context& c, uint32_t n_jobs, bool tagged, int levels,
                      const BuildOut* out = nullptr);

// Streamed build (svo_stream.hip): begin (emits the leading data records), one call per
// partition after its voxels were added (builds, emits and empties the voxel stream),
// end (levels above the partitions, the root; the counts). Without a sink the records
// collect in StreamState::out_nodes / out_data and end moves them to nodes / records.
hipError_t stream_begin(Context& c, uint64_t gridsize, uint64_t n_parts, int levels,
                        SinkFn sink, void* user);
hipError_t stream_partition(Context& c, uint64_t part);
hipError_t stream_end(Context& c, uint64_t* n_nodes, uint64_t* n_data, int* hazard);
// Free the streamed build's output buffers and pinned chunks.
void stream_release(Context& c);

hipError_t fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
hipError_t fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s);

void set_error(const std::string& msg);

}  // namespace svo
