// Library-internal state and stage drivers behind the C ABI (include/svo_amd.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

namespace svo {

// Sparse occupancy grid: the partition's Morton range is cut into aligned bricks of
// 2^kBrickBits voxels (16^3). A dense table maps brick -> pool slot (0 = empty). A slot
// holds the brick's voxels as bytes (marked by plain stores) and, after compaction, as
// 64 bit words, both in Morton order.
constexpr int kBrickBits = 12;
constexpr int kBrickVox = 1 << kBrickBits;
constexpr int kBrickWords = 1 << (kBrickBits - 6);

// Device counters block (one 256-byte line of u64s).
enum Counter : int {
  kCtrPoolCount = 0,  // bricks allocated from the pool
  kCtrOverflow = 1,   // pool exhausted during a pass
  kCtrLargeCount = 2, // triangles deferred to the column-parallel tier
  kCtrNumOcc = 3,     // occupied bricks after compaction
  kCtrNumVox = 4,     // voxels of the current partition
  kCtrLevelN = 5,     // node count of the level being built
  kCtrCols = 6,       // columns of the large tier
  kCtrError = 7,      // sticky: internal consistency failures (cleared per job)
  kNumCounters = 32
};

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
  int small_cols = 1024;
  int block = 256;
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

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  VoxTuning tune;

  // mesh (device-resident records, 9 or 21 floats each)
  DevBuf tris;
  uint64_t n_tris = 0;
  int payload = 0;

  // partitioning
  DevBuf hitbox, part_lists;
  std::vector<uint64_t> part_off, part_cnt;
  uint64_t n_parts = 0;

  // voxelizer scratch (kept zeroed between partitions: table, pool)
  DevBuf table, pool, pool8;  // pool: bit words per slot; pool8: bytes per slot
  uint64_t table_entries = 0;
  uint64_t pool_cap = 0;  // bricks
  DevBuf counters;
  DevBuf large_tri, large_cols, large_off;
  DevBuf occ_list, bcount, wprefix, winner;
  DevBuf scan_scratch;

  // accumulated voxel stream across partitions (sorted by construction)
  DevBuf codes;    // u64 Morton codes
  DevBuf records;  // payload: 32 B records, slot 0 reserved for the null record
  uint64_t n_vox = 0;
  uint64_t nfilled = 0;

  // build output
  DevBuf nodes, data;
  uint64_t n_nodes = 0, n_data = 0;
  int hazard = 0;
  DevBuf lvl_code, lvl_size, lvl_mask, lvl_first, lvl_start, lvl_aux, lvl_leafp;

  StageTimes times;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  bool ktiming = false;
  std::vector<PendingTimer> kpending;
  std::vector<hipEvent_t> evpool;
  std::map<std::string, KernelStat> kstats;
};

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

// Stage drivers (return hipSuccess or an error; message in last error).
hipError_t partition_mesh(Context& c, float unitlength, uint64_t gridsize, uint64_t n_parts);
hipError_t voxelize_partition(Context& c, uint64_t part, uint64_t morton_start,
                              uint64_t morton_end, float unitlength, uint64_t gridsize,
                              int color_mode, uint64_t* nfilled);
hipError_t build_octree(Context& c, uint64_t gridsize, int levels);

hipError_t fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
hipError_t fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s);

void set_error(const std::string& msg);

}  // namespace svo
