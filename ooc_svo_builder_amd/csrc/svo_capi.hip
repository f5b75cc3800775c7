// C ABI (include/svo_amd.h) over the stage drivers, and the in-memory main.cpp flow.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>

#include "../../include/svo_amd.h"
#include "svo_internal.h"

struct svo_ctx {
  svo::Context c;
};

namespace svo {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

// Growth policy: a buffer that has to grow takes 1.5x its old size when that is more
// (bounded by twice the request), so that a sequence of slightly larger requests (the
// batches of a server) settles after a few reallocations; every hipFree waits for the
// whole device.
static size_t grown(const DevBuf& b, size_t bytes) {
  if (!b.p) return std::max<size_t>(bytes, 256);
  return std::max<size_t>(bytes, std::min<size_t>(b.cap + b.cap / 2, 2 * bytes));
}

hipError_t ensure(DevBuf& b, size_t bytes) {
  if (bytes <= b.cap && b.p) return hipSuccess;
  const size_t want = grown(b, bytes);
  if (b.p) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
  }
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    set_error("hipMalloc of " + std::to_string(want) + " bytes failed");
    return hipErrorOutOfMemory;
  }
  b.cap = want;
  return hipSuccess;
}

hipError_t ensure_keep(DevBuf& b, size_t bytes, size_t used, hipStream_t s) {
  if (bytes <= b.cap && b.p) return hipSuccess;
  const size_t want = grown(b, bytes);
  void* np = nullptr;
  hipError_t e = hipMalloc(&np, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error("hipMalloc of " + std::to_string(want) + " bytes failed");
    return hipErrorOutOfMemory;
  }
  if (b.p) {
    if (used) {
      e = hipMemcpyAsync(np, b.p, std::min(used, b.cap), hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return e;
    }
    e = stream_sync(s);
    if (e != hipSuccess) return e;
    (void)hipFree(b.p);
  }
  b.p = np;
  b.cap = want;
  return hipSuccess;
}

void release(DevBuf& b) {
  if (b.p) {
    (void)hipDeviceSynchronize();
    (void)hipFree(b.p);
  }
  b.p = nullptr;
  b.cap = 0;
}

static hipEvent_t take_event(Context& c) {
  if (!c.evpool.empty()) {
    hipEvent_t e = c.evpool.back();
    c.evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void ktimer_begin(Context& c, const char* name) {
  if (!c.ktiming) return;
  PendingTimer t;
  t.name = name;
  t.a = take_event(c);
  t.b = take_event(c);
  if (!t.a || !t.b) return;
  (void)hipEventRecord(t.a, c.stream);
  c.kpending.push_back(t);
}

void ktimer_end(Context& c) {
  if (!c.ktiming || c.kpending.empty()) return;
  (void)hipEventRecord(c.kpending.back().b, c.stream);
}

void ktimer_flush(Context& c) {
  for (PendingTimer& t : c.kpending) {
    float ms = 0;
    if (event_sync(t.b) == hipSuccess && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      KernelStat& k = c.kstats[t.name];
      k.launches++;
      k.ms += ms;
    }
    c.evpool.push_back(t.a);
    c.evpool.push_back(t.b);
  }
  c.kpending.clear();
}

}  // namespace svo

using namespace svo;

static int fail(hipError_t e, const char* where) {
  if (e == hipErrorUnknown && !g_err.empty() && g_err.rfind("the output sink", 0) == 0)
    return SVO_ERR_IO;
  if (e == hipErrorOutOfMemory) {
    if (g_err.empty()) g_err = std::string(where) + ": out of device memory";
    return SVO_ERR_NOMEM;
  }
  if (e == hipErrorNotSupported) return SVO_ERR_UNSUPPORTED;
  if (e == hipErrorInvalidValue) {
    if (g_err.empty()) g_err = std::string(where) + ": invalid value";
    return SVO_ERR_INVALID;
  }
  std::string m = std::string(where) + ": " + hipGetErrorString(e);
  if (!g_err.empty()) m += " (" + g_err + ")";
  g_err = m;
  return SVO_ERR_DEVICE;
}

#define CHECK_CTX()                         \
  do {                                      \
    if (!ctx) {                             \
      g_err = "null context";               \
      return SVO_ERR_INVALID;               \
    }                                       \
    g_err.clear();                          \
    hipError_t _d = hipSetDevice(ctx->c.device); \
    if (_d != hipSuccess) return fail(_d, "hipSetDevice"); \
  } while (0)

#define TRYC(x, where)                          \
  do {                                          \
    hipError_t _e = (x);                        \
    if (_e != hipSuccess) return fail(_e, where); \
  } while (0)

static int is_pow2_u64(uint64_t x) { return x != 0 && !(x & (x - 1)); }

// What the sticky consistency-failure counter counts (kCtrError): a look-back wait of a
// single-pass scan that ran out of spins, a brick without a pool slot on the touched-brick
// path, and a payload voxel whose covering triangle was not found.
static const char* const kErrSources =
    " consistency failures (scan look-back time-outs, brick pool slots not found or payload "
    "voxels without a covering triangle)";

// the sticky consistency-failure counter
static hipError_t read_err(Context& c, unsigned long long* err) {
  *err = c.err_count;  // read with the build's results
  return hipSuccess;
}

// Copies on the context's own stream (never the null stream, which would wait for every
// other context's work), direction from the pointers.
static hipError_t copy_sync(Context& c, void* dst, const void* src, size_t bytes) {
  if (!bytes) return hipSuccess;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, c.stream);
  if (e != hipSuccess) return e;
  return stream_sync(c.stream);
}
static int ctz64(uint64_t x) { return x ? __builtin_ctzll(x) : 64; }

// The plain flow's single job: the mesh loaded by svo_mesh_load, an empty voxel stream.
static void reset_plain(Context& c) {
  c.jobs.assign(1, Job{});
  Job& J = c.jobs[0];
  J.tris = c.tris.as<float>();
  J.n_tris = c.n_tris;
  J.payload = c.payload;
  c.part_used = 0;
  c.n_vox = 0;
  c.nfilled = 0;
  c.n_nodes = c.n_data = 0;
  c.hazard = 0;
  c.batch_state = false;
  c.data_in_records2 = false;
  c.n_bricks = 0;
  c.bricks_ok = true;
  c.codes_lazy = false;
  c.st.active = false;  // a new mesh ends any streamed build in progress
}

// The 6-significant-digit text round trip a bbox value takes through the .trip header
// (trip_tools.h:112 writes it, main.cpp:306-307 re-parses it).
static float text_roundtrip(float v) {
  std::ostringstream o;
  o << v;
  std::istringstream i(o.str());
  float r = 0.0f;
  i >> r;
  return r;
}

extern "C" {

const char* svo_last_error(void) { return g_err.c_str(); }
const char* svo_version(void) { return "ooc_svo_builder_amd 0.1 (svo_builder 1.6.4 formats)"; }

int svo_ctx_create(int device, svo_ctx** out) {
  g_err.clear();
  if (!out) {
    g_err = "null output pointer";
    return SVO_ERR_INVALID;
  }
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    g_err = "no HIP device available";
    return SVO_ERR_DEVICE;
  }
  if (device < 0 || device >= ndev) {
    g_err = "device index out of range";
    return SVO_ERR_INVALID;
  }
  svo_ctx* ctx = new (std::nothrow) svo_ctx();
  if (!ctx) {
    g_err = "host allocation failed";
    return SVO_ERR_NOMEM;
  }
  ctx->c.device = device;
  if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreate(&ctx->c.stream)) != hipSuccess) {
    delete ctx;
    return fail(e, "svo_ctx_create");
  }
  ctx->c.own_stream = true;
  if ((e = ensure(ctx->c.counters, kNumCounters * 8)) != hipSuccess ||
      (e = hipEventCreate(&ctx->c.ev0)) != hipSuccess ||
      (e = hipEventCreate(&ctx->c.ev1)) != hipSuccess ||
      (e = hipEventCreate(&ctx->c.ev2)) != hipSuccess ||
      (e = hipEventCreate(&ctx->c.ev3)) != hipSuccess) {
    svo_ctx_destroy(ctx);
    return fail(e, "svo_ctx_create");
  }
  {
    // dense voxel grids up to half the free HBM (at most 128 GiB): 288 GB of HBM3E holds
    // the 2048^3 grid (8 GiB), a 4096^3 partition (64 GiB), or a batch's grids in few
    // passes (the 512-mesh mix: 4 passes at 64 GiB, 2 at 128 GiB, 4% faster) next to
    // everything else; the grid is allocated as passes need it
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
      ctx->c.dense_limit = std::min<uint64_t>((uint64_t)free_b / 2, 128ull << 30);
    const char* env = getenv("SVO_DENSE_LIMIT_MB");
    if (env) ctx->c.dense_limit = (uint64_t)strtoull(env, nullptr, 10) << 20;
  }
  reset_plain(ctx->c);
  *out = ctx;
  return SVO_OK;
}

void svo_ctx_destroy(svo_ctx* ctx) {
  if (!ctx) return;
  Context& c = ctx->c;
  (void)hipSetDevice(c.device);
  if (c.stream) (void)stream_sync(c.stream);
  for (hipGraphExec_t x : c.gexec)
    if (x) (void)hipGraphExecDestroy(x);
  c.gexec.clear();
  c.glog.clear();
  DevBuf* bufs[] = {&c.tris, &c.jobdev, &c.batch_tris, &c.hitbox, &c.part_lists, &c.unitdev,
                    &c.table, &c.pool, &c.pool8, &c.grid8, &c.counters, &c.large_unit,
                    &c.large_tri, &c.large_cols, &c.large_off, &c.occ_list, &c.bcount,
                    &c.wprefix, &c.winner, &c.scan_scratch, &c.codes, &c.records, &c.nodes,
                    &c.lvl_code, &c.lvl_size, &c.lvl_mask, &c.lvl_first, &c.lvl_start,
                    &c.lvl_aux, &c.lvl_leafp, &c.lvl_pidx, &c.lvl_isize, &c.lvl_cache,
                    &c.lvl_pstart, &c.records2, &c.bwords, &c.bcode, &c.bvoff,
                    &c.scan_status, &c.lvl_sib, &c.up_tiles, &c.bshape, &c.large_walk,
                    &c.payload_map};
  for (DevBuf* b : bufs) release(*b);
  stream_release(c);
  if (c.aux) svo_ctx_destroy((svo_ctx*)c.aux);
  c.aux = nullptr;
  if (c.stage) (void)hipHostFree(c.stage);
  c.stage = nullptr;
  if (c.hread) (void)hipHostFree(c.hread);
  c.hread = nullptr;
  for (int k = 0; k < 2; k++) {
    if (c.ingest[k]) (void)hipHostFree(c.ingest[k]);
    if (c.ingest_ev[k]) (void)hipEventDestroy(c.ingest_ev[k]);
  }
  ktimer_flush(c);
  for (hipEvent_t e : c.evpool) (void)hipEventDestroy(e);
  for (hipEvent_t ev : {c.ev0, c.ev1, c.ev2, c.ev3})
    if (ev) (void)hipEventDestroy(ev);
  if (c.own_stream && c.stream) (void)hipStreamDestroy(c.stream);
  delete ctx;
}

int svo_ctx_set_stream(svo_ctx* ctx, void* hip_stream) {
  CHECK_CTX();
  Context& c = ctx->c;
  TRYC(stream_sync(c.stream), "svo_ctx_set_stream");
  if (hip_stream) {
    if (c.own_stream) (void)hipStreamDestroy(c.stream);
    c.own_stream = false;
    c.stream = (hipStream_t)hip_stream;
  } else if (!c.own_stream) {
    TRYC(hipStreamCreate(&c.stream), "svo_ctx_set_stream");
    c.own_stream = true;
  }
  return SVO_OK;
}

void* svo_ctx_stream(svo_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

int svo_ctx_set_tuning(svo_ctx* ctx, int small_cols, uint64_t dense_limit_mb) {
  CHECK_CTX();
  if (small_cols > 0) ctx->c.tune.small_cols = small_cols;
  if (dense_limit_mb != UINT64_MAX) ctx->c.dense_limit = dense_limit_mb << 20;
  return SVO_OK;
}

int svo_ctx_set_graphs(svo_ctx* ctx, int on) {
  CHECK_CTX();
  ctx->c.graph_policy = on ? 1 : 0;
  return SVO_OK;
}

uint64_t svo_estimate_partitions(uint64_t gridsize, uint64_t memory_limit_mb) {
  const uint64_t required = (gridsize * gridsize * gridsize) / 1024 / 1024;
  if (required <= memory_limit_mb) return 1;
  uint64_t np = 1, req = required;
  while (req > memory_limit_mb) {
    req /= 8;
    np *= 8;
  }
  return np;
}

int svo_mesh_load(svo_ctx* ctx, const float* records, uint64_t n_triangles, int payload,
                  int src_on_device) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (n_triangles && !records) {
    g_err = "null records";
    return SVO_ERR_INVALID;
  }
  const size_t bytes = n_triangles * (payload ? 21 : 9) * sizeof(float);
  TRYC(ensure(c.tris, bytes), "svo_mesh_load");
  if (bytes)
    TRYC(hipMemcpyAsync(c.tris.p, records, bytes,
                        src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c.stream),
         "svo_mesh_load");
  TRYC(fill_u64(c.counters.as<uint64_t>(), 0, kNumCounters, c.stream), "svo_mesh_load");
  TRYC(stream_sync(c.stream), "svo_mesh_load");
  c.n_tris = n_triangles;
  c.payload = payload ? 1 : 0;
  reset_plain(c);
  return SVO_OK;
}

int svo_mesh_load_file(svo_ctx* ctx, const char* tridata_path, uint64_t n_triangles,
                       int payload, uint64_t* got_out) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (got_out) *got_out = 0;
  if (!tridata_path) {
    g_err = "null path";
    return SVO_ERR_INVALID;
  }
  FILE* f = fopen(tridata_path, "rb");
  if (!f) {
    g_err = std::string("cannot open ") + tridata_path;
    return SVO_ERR_INVALID;
  }
  const uint64_t R = payload ? 84 : 36;
  const uint64_t bytes = n_triangles * R;
  hipError_t e = ensure(c.tris, bytes);
  if (e != hipSuccess) {
    fclose(f);
    return fail(e, "svo_mesh_load_file");
  }
  // two pinned chunks: the file read of one overlaps the copy of the other
  const size_t chunk = 32u << 20;
  if (c.ingest_cap < chunk) {
    for (int k = 0; k < 2; k++) {
      if (c.ingest[k]) (void)hipHostFree(c.ingest[k]);
      c.ingest[k] = nullptr;
      if ((e = hipHostMalloc(&c.ingest[k], chunk, hipHostMallocDefault)) != hipSuccess ||
          (!c.ingest_ev[k] && (e = hipEventCreate(&c.ingest_ev[k])) != hipSuccess)) {
        fclose(f);
        c.ingest_cap = 0;
        return fail(e, "svo_mesh_load_file");
      }
    }
    c.ingest_cap = chunk;
  }
  uint64_t off = 0;
  for (uint64_t k = 0; off < bytes; k++) {
    const int b = (int)(k & 1);
    if (k >= 2 && (e = event_sync(c.ingest_ev[b])) != hipSuccess) break;
    const size_t want = (size_t)std::min<uint64_t>(chunk, bytes - off);
    const size_t got = fread(c.ingest[b], 1, want, f);
    if (got &&
        ((e = hipMemcpyAsync((char*)c.tris.p + off, c.ingest[b], got, hipMemcpyHostToDevice,
                             c.stream)) != hipSuccess ||
         (e = hipEventRecord(c.ingest_ev[b], c.stream)) != hipSuccess))
      break;
    off += got;
    if (got < want) break;
  }
  fclose(f);
  if (e != hipSuccess) return fail(e, "svo_mesh_load_file");
  TRYC(fill_u64(c.counters.as<uint64_t>(), 0, kNumCounters, c.stream), "svo_mesh_load_file");
  TRYC(stream_sync(c.stream), "svo_mesh_load_file");
  if (got_out) *got_out = off / R;
  if (off < bytes) {
    g_err = std::string(tridata_path) + " holds " + std::to_string(off / R) +
            " triangles, the header says " + std::to_string(n_triangles);
    return SVO_ERR_INVALID;
  }
  c.n_tris = n_triangles;
  c.payload = payload ? 1 : 0;
  reset_plain(c);
  return SVO_OK;
}

int svo_partition(svo_ctx* ctx, float unitlength, uint64_t gridsize, uint64_t n_partitions,
                  uint64_t* counts_out) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (c.batch_state) reset_plain(c);
  if (!is_pow2_u64(gridsize) || n_partitions == 0) {
    g_err = "gridsize must be a power of 2 and n_partitions > 0";
    return SVO_ERR_INVALID;
  }
  uint64_t p = n_partitions;
  while (p % 8 == 0) p /= 8;
  if (p != 1 || n_partitions > gridsize * gridsize * gridsize) {
    g_err = "n_partitions must be a power of 8 not above gridsize^3";
    return SVO_ERR_INVALID;
  }
  c.part_used = 0;
  TRYC(partition_job(c, 0, unitlength, gridsize, n_partitions), "svo_partition");
  if (counts_out)
    for (uint64_t j = 0; j < n_partitions; j++) counts_out[j] = c.jobs[0].part_cnt[j];
  return SVO_OK;
}

int svo_partition_indices(svo_ctx* ctx, uint64_t partition, uint32_t* host_out) {
  CHECK_CTX();
  Context& c = ctx->c;
  const Job& J = c.jobs[0];
  if (partition >= J.n_parts) {
    g_err = "partition index out of range";
    return SVO_ERR_INVALID;
  }
  if (J.n_parts == 1) {
    for (uint64_t i = 0; i < c.n_tris; i++) host_out[i] = (uint32_t)i;
    return SVO_OK;
  }
  if (J.part_cnt[partition])
    TRYC(copy_sync(c, host_out, c.part_lists.as<uint32_t>() + J.part_off[partition],
                   J.part_cnt[partition] * 4),
         "svo_partition_indices");
  return SVO_OK;
}

int svo_write_trip(svo_ctx* ctx, const char* base, uint64_t gridsize, const float bbox[6],
                   int version) {
  CHECK_CTX();
  Context& c = ctx->c;
  const Job& J = c.jobs[0];
  if (!base || !bbox || c.batch_state || J.n_parts == 0 || J.part_cnt.size() != J.n_parts) {
    g_err = "svo_write_trip: partition the loaded mesh first (svo_partition)";
    return SVO_ERR_INVALID;
  }
  const int stride = c.payload ? 21 : 9;
  const uint64_t rec = (uint64_t)stride * 4;
  // partition i's records in mesh order, through a device staging buffer and a pinned
  // host buffer, 1 Mi triangles at a time (BBoxBuffer.h:56-67 opens a partition's file at
  // its first flush, so empty partitions get none)
  const uint64_t chunk = 1u << 20;
  DevBuf dstage;
  void* hstage = nullptr;
  hipError_t e = ensure(dstage, chunk * rec);
  if (e == hipSuccess) e = hipHostMalloc(&hstage, chunk * rec, hipHostMallocDefault);
  if (e != hipSuccess) {
    release(dstage);
    return fail(e, "svo_write_trip");
  }
  int rc = SVO_OK;
  for (uint64_t i = 0; i < J.n_parts && rc == SVO_OK; i++) {
    const uint64_t n = J.part_cnt[i];
    if (!n) continue;
    const std::string fn = std::string(base) + "_" + std::to_string(i) + ".tripdata";
    FILE* f = fopen(fn.c_str(), "wb");
    if (!f) {
      g_err = "svo_write_trip: cannot open " + fn;
      rc = SVO_ERR_IO;
      break;
    }
    const uint32_t* list = J.n_parts > 1 ? c.part_lists.as<uint32_t>() + J.part_off[i] : nullptr;
    for (uint64_t k = 0; k < n && rc == SVO_OK; k += chunk) {
      const uint64_t m = std::min(chunk, n - k);
      e = list ? gather_records(c, c.tris.as<float>(), stride, list + k, m, dstage.as<float>())
               : hipMemcpyAsync(dstage.p, c.tris.as<char>() + k * rec, m * rec,
                                hipMemcpyDeviceToDevice, c.stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(hstage, dstage.p, m * rec, hipMemcpyDeviceToHost, c.stream);
      if (e == hipSuccess) e = stream_sync(c.stream);
      if (e != hipSuccess) {
        rc = fail(e, "svo_write_trip");
      } else if (fwrite(hstage, 1, m * rec, f) != m * rec) {
        g_err = "svo_write_trip: writing " + fn + " failed";
        rc = SVO_ERR_IO;
      }
    }
    if (fclose(f) != 0 && rc == SVO_OK) {
      g_err = "svo_write_trip: writing " + fn + " failed";
      rc = SVO_ERR_IO;
    }
  }
  release(dstage);
  (void)hipHostFree(hstage);
  if (rc != SVO_OK) return rc;
  // trip_tools.h:104-125 (floats in the default 6 significant digits)
  std::ofstream o((std::string(base) + ".trip").c_str());
  o << "#trip " << version << std::endl;
  o << "gridsize " << gridsize << std::endl;
  o << "n_triangles " << c.n_tris << std::endl;
  o << "bbox  " << bbox[0] << " " << bbox[1] << " " << bbox[2] << " " << bbox[3] << " "
    << bbox[4] << " " << bbox[5] << std::endl;
  o << "geo_only " << (c.payload ? 0 : 1) << std::endl;
  o << "n_partitions " << J.n_parts << std::endl;
  for (uint64_t i = 0; i < J.n_parts; i++) o << i << " " << J.part_cnt[i] << std::endl;
  o << "END" << std::endl;
  o.close();
  if (!o) {
    g_err = std::string("svo_write_trip: writing ") + base + ".trip failed";
    return SVO_ERR_IO;
  }
  return SVO_OK;
}

int svo_voxelize(svo_ctx* ctx, uint64_t partition, uint64_t morton_start, uint64_t morton_end,
                 float unitlength, uint64_t gridsize, int color_mode, uint64_t* nfilled_out) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (c.batch_state) reset_plain(c);
  if (!is_pow2_u64(gridsize) || gridsize > (1u << 15) || morton_end <= morton_start ||
      morton_end > gridsize * gridsize * gridsize || color_mode < 0 || color_mode > 3) {
    g_err = "invalid voxelize arguments";
    return SVO_ERR_INVALID;
  }
  // the reference's partitions are aligned Morton cubes (main.cpp:329-333)
  const uint64_t span = morton_end - morton_start;
  if (!is_pow2_u64(span) || (ctz64(span) % 3) != 0 || morton_start % span != 0) {
    g_err = "the Morton range must be an aligned cube (a partition)";
    return SVO_ERR_INVALID;
  }
  Job& J = c.jobs[0];
  if (J.n_parts > 1 && partition >= J.n_parts) {
    g_err = "partition index out of range";
    return SVO_ERR_INVALID;
  }
  if (J.n_parts <= 1) partition = 0;
  J.gridsize = gridsize;
  J.unit_vox = unitlength;
  J.color_mode = color_mode;
  VoxUnit u;
  u.job = 0;
  u.tris = c.tris.as<float>();
  u.list = (J.n_parts > 1) ? c.part_lists.as<uint32_t>() + J.part_off[partition] : nullptr;
  u.n = (J.n_parts > 1) ? J.part_cnt[partition] : c.n_tris;
  u.stride = c.payload ? 21 : 9;
  u.mstart = morton_start;
  u.mend = morton_end;
  u.unit = unitlength;
  u.gridsize = gridsize;
  uint64_t nf = 0;
  TRYC(voxelize_units(c, std::vector<VoxUnit>{u}, false, &nf), "svo_voxelize");
  c.nfilled += nf;
  if (nfilled_out) *nfilled_out = nf;
  return SVO_OK;
}

int svo_add_voxels(svo_ctx* ctx, const uint64_t* codes, const void* records32, uint64_t n) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (c.batch_state) reset_plain(c);
  if (n == 0) return SVO_OK;
  if (!codes || (c.payload && !records32)) {
    g_err = "null codes or records";
    return SVO_ERR_INVALID;
  }
  uint64_t last = 0;
  TRYC(ensure_codes(c), "svo_add_voxels");
  if (c.n_vox)
    TRYC(copy_sync(c, &last, c.codes.as<uint64_t>() + c.n_vox - 1, 8),
         "svo_add_voxels");
  for (uint64_t i = 0; i < n; i++) {
    if ((c.n_vox || i) && codes[i] <= last) {
      g_err = "voxel codes must be strictly increasing";
      return SVO_ERR_INVALID;
    }
    last = codes[i];
  }
  c.bricks_ok = false;  // these voxels come without their bricks
  TRYC(ensure_keep(c.codes, (c.n_vox + n) * 8, c.n_vox * 8, c.stream), "svo_add_voxels");
  TRYC(hipMemcpyAsync(c.codes.as<uint64_t>() + c.n_vox, codes, n * 8, hipMemcpyHostToDevice,
                      c.stream),
       "svo_add_voxels");
  if (c.payload) {
    TRYC(ensure_keep(c.records, (1 + c.n_vox + n) * 32, (1 + c.n_vox) * 32, c.stream),
         "svo_add_voxels");
    TRYC(hipMemcpyAsync((char*)c.records.p + (1 + c.n_vox) * 32, records32, n * 32,
                        hipMemcpyHostToDevice, c.stream),
         "svo_add_voxels");
  }
  TRYC(stream_sync(c.stream), "svo_add_voxels");
  c.n_vox += n;
  return SVO_OK;
}

int svo_build(svo_ctx* ctx, uint64_t gridsize, int generate_levels, uint64_t* n_nodes_out,
              uint64_t* n_data_out, int* hazard_out) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (!is_pow2_u64(gridsize)) {
    g_err = "gridsize must be a power of 2";
    return SVO_ERR_INVALID;
  }
  if (c.batch_state) reset_plain(c);
  Job& J = c.jobs[0];
  J.gridsize = gridsize;
  J.voff = 0;
  J.nvox = c.n_vox;
  hipError_t e = build_jobs(c, 1, false, generate_levels);
  if (e != hipSuccess) return fail(e, "svo_build");
  c.hazard = J.hazard;
  unsigned long long err = 0;
  TRYC(read_err(c, &err),
       "svo_build");
  if (err) {
    g_err = "internal: " + std::to_string(err) + kErrSources;
    return SVO_ERR_DEVICE;
  }
  if (n_nodes_out) *n_nodes_out = c.n_nodes;
  if (n_data_out) *n_data_out = c.n_data;
  if (hazard_out) *hazard_out = c.hazard;
  return SVO_OK;
}

int svo_copy_nodes(svo_ctx* ctx, void* host_dst, uint64_t first, uint64_t count) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (first + count > c.n_nodes) {
    g_err = "node range out of bounds";
    return SVO_ERR_INVALID;
  }
  if (count)
    TRYC(copy_sync(c, host_dst, (const char*)c.nodes.p + first * 24, count * 24),
         "svo_copy_nodes");
  return SVO_OK;
}

int svo_copy_data(svo_ctx* ctx, void* host_dst, uint64_t first, uint64_t count) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (first + count > c.n_data) {
    g_err = "data range out of bounds";
    return SVO_ERR_INVALID;
  }
  if (count)
    TRYC(copy_sync(c, host_dst, (const char*)c.data_out().p + first * 32, count * 32),
         "svo_copy_data");
  return SVO_OK;
}

const void* svo_nodes_device(svo_ctx* ctx) { return ctx ? ctx->c.nodes.p : nullptr; }
const void* svo_data_device(svo_ctx* ctx) { return ctx ? ctx->c.data_out().p : nullptr; }

uint64_t svo_voxel_count(svo_ctx* ctx) { return ctx ? ctx->c.n_vox : 0; }

int svo_copy_codes(svo_ctx* ctx, uint64_t* host_dst, uint64_t first, uint64_t count) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (first + count > c.n_vox) {
    g_err = "code range out of bounds";
    return SVO_ERR_INVALID;
  }
  TRYC(ensure_codes(c), "svo_copy_codes");
  if (count)
    TRYC(copy_sync(c, host_dst, c.codes.as<uint64_t>() + first, count * 8),
         "svo_copy_codes");
  return SVO_OK;
}

int svo_ctx_set_kernel_timing(svo_ctx* ctx, int on) {
  CHECK_CTX();
  TRYC(stream_sync(ctx->c.stream), "svo_ctx_set_kernel_timing");
  ktimer_flush(ctx->c);
  ctx->c.ktiming = on != 0;
  return SVO_OK;
}

int svo_kernel_timing(svo_ctx* ctx, const char* name, uint64_t* launches, double* total_ms) {
  CHECK_CTX();
  TRYC(stream_sync(ctx->c.stream), "svo_kernel_timing");
  ktimer_flush(ctx->c);
  auto it = ctx->c.kstats.find(name ? name : "");
  if (launches) *launches = it == ctx->c.kstats.end() ? 0 : it->second.launches;
  if (total_ms) *total_ms = it == ctx->c.kstats.end() ? 0.0 : it->second.ms;
  return SVO_OK;
}

int svo_kernel_timing_reset(svo_ctx* ctx) {
  CHECK_CTX();
  TRYC(stream_sync(ctx->c.stream), "svo_kernel_timing_reset");
  ktimer_flush(ctx->c);
  ctx->c.kstats.clear();
  return SVO_OK;
}

int svo_stream_begin(svo_ctx* ctx, uint64_t gridsize, uint64_t n_partitions,
                     int generate_levels, svo_sink_fn sink, void* user) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (c.batch_state) {
    g_err = "no mesh loaded (the context holds a batch)";
    return SVO_ERR_INVALID;
  }
  if (!is_pow2_u64(gridsize) || gridsize > (1u << 15) || n_partitions == 0) {
    g_err = "gridsize must be a power of 2 up to 32768 and n_partitions > 0";
    return SVO_ERR_INVALID;
  }
  uint64_t p = n_partitions;
  while (p % 8 == 0) p /= 8;
  if (p != 1 || n_partitions > gridsize * gridsize * gridsize) {
    g_err = "n_partitions must be a power of 8 not above gridsize^3";
    return SVO_ERR_INVALID;
  }
  TRYC(stream_begin(c, gridsize, n_partitions, generate_levels, (SinkFn)sink, user),
       "svo_stream_begin");
  return SVO_OK;
}

int svo_stream_partition(svo_ctx* ctx, uint64_t partition) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (!c.st.active) {
    g_err = "no streamed build in progress (svo_stream_begin)";
    return SVO_ERR_INVALID;
  }
  if (partition < c.st.next_part || partition >= c.st.n_parts) {
    g_err = "partitions must come in increasing order, below n_partitions";
    return SVO_ERR_INVALID;
  }
  hipError_t e = stream_partition(c, partition);
  if (e != hipSuccess) {
    c.st.active = false;
    return fail(e, "svo_stream_partition");
  }
  return SVO_OK;
}

int svo_stream_end(svo_ctx* ctx, uint64_t* n_nodes_out, uint64_t* n_data_out, int* hazard_out) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (!c.st.active) {
    g_err = "no streamed build in progress (svo_stream_begin)";
    return SVO_ERR_INVALID;
  }
  hipError_t e = stream_end(c, n_nodes_out, n_data_out, hazard_out);
  if (e != hipSuccess) {
    c.st.active = false;
    return fail(e, "svo_stream_end");
  }
  return SVO_OK;
}

// main.cpp:281-399 in memory. One partition: voxelize and build in HBM (svo_build).
// Several partitions, or a sink: the streamed build, one partition's subtree at a time
// (records to the sink, or collected in HBM when sink == NULL).
static int run_impl(svo_ctx* ctx, const float* records, uint64_t n_triangles, int payload,
                    int src_on_device, const float bbox[6], uint64_t gridsize,
                    uint64_t memory_limit_mb, int color_mode, int generate_levels,
                    bool streamed, svo_sink_fn sink, void* user, svo_run_info* info) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (!bbox || !is_pow2_u64(gridsize) || memory_limit_mb <= 1) {
    g_err = "invalid run arguments";
    return SVO_ERR_INVALID;
  }
  const uint64_t l0 = t_launches, s0 = t_syncs;
  int r = SVO_OK;
  if (records) {
    r = svo_mesh_load(ctx, records, n_triangles, payload, src_on_device);
    if (r) return r;
  } else {  // the mesh already loaded: reset the stream and outputs only
    if ((payload ? 1 : 0) != c.payload) {
      g_err = "payload flag differs from the loaded mesh";
      return SVO_ERR_INVALID;
    }
    reset_plain(c);
    TRYC(fill_u64(c.counters.as<uint64_t>(), 0, kNumCounters, c.stream), "svo_run");
  }
  const uint64_t np = svo_estimate_partitions(gridsize, memory_limit_mb);
  streamed = streamed || np > 1;
  const float u_part = (bbox[3] - bbox[0]) / (float)gridsize;  // partitioner.cpp:45
  const float u_vox =
      (text_roundtrip(bbox[3]) - text_roundtrip(bbox[0])) / (float)gridsize;  // main.cpp:311
  // stage boundaries as events, read once the build has synchronized
  float ms_p = 0, ms_v = 0, ms_b = 0;
  TRYC(hipEventRecord(c.ev0, c.stream), "svo_run");
  std::vector<uint64_t> counts(np, 0);
  r = svo_partition(ctx, u_part, gridsize, np, counts.data());
  if (r) return r;
  TRYC(hipEventRecord(c.ev1, c.stream), "svo_run");
  const uint64_t morton_part = (gridsize * gridsize * gridsize) / np;
  uint64_t nn = 0, nd = 0;
  int hz = 0;
  if (streamed) {
    TRYC(event_sync(c.ev1), "svo_run");
    TRYC(hipEventElapsedTime(&ms_p, c.ev0, c.ev1), "svo_run");
    r = svo_stream_begin(ctx, gridsize, np, generate_levels, sink, user);
    if (r) return r;
    for (uint64_t i = 0; i < np; i++) {
      if (counts[i] == 0) continue;  // main.cpp:330
      TRYC(hipEventRecord(c.ev0, c.stream), "svo_run");
      r = svo_voxelize(ctx, i, i * morton_part, (i + 1) * morton_part, u_vox, gridsize,
                       color_mode, nullptr);
      if (r) return r;
      TRYC(hipEventRecord(c.ev1, c.stream), "svo_run");
      r = svo_stream_partition(ctx, i);
      if (r) return r;
      TRYC(hipEventRecord(c.ev2, c.stream), "svo_run");
      TRYC(event_sync(c.ev2), "svo_run");
      float a = 0, b = 0;
      TRYC(hipEventElapsedTime(&a, c.ev0, c.ev1), "svo_run");
      TRYC(hipEventElapsedTime(&b, c.ev1, c.ev2), "svo_run");
      ms_v += a;
      ms_b += b;
    }
    TRYC(hipEventRecord(c.ev0, c.stream), "svo_run");
    r = svo_stream_end(ctx, &nn, &nd, &hz);
    if (r) return r;
    TRYC(hipEventRecord(c.ev1, c.stream), "svo_run");
    TRYC(event_sync(c.ev1), "svo_run");
    float b = 0;
    TRYC(hipEventElapsedTime(&b, c.ev0, c.ev1), "svo_run");
    ms_b += b;
  } else {
    for (uint64_t i = 0; i < np; i++) {
      if (counts[i] == 0) continue;  // main.cpp:330
      r = svo_voxelize(ctx, i, i * morton_part, (i + 1) * morton_part, u_vox, gridsize,
                       color_mode, nullptr);
      if (r) return r;
    }
    TRYC(hipEventRecord(c.ev2, c.stream), "svo_run");
    c.build_done_ev = c.ev3;  // recorded by the build before its final read-back
    r = svo_build(ctx, gridsize, generate_levels, &nn, &nd, &hz);
    c.build_done_ev = nullptr;
    if (r) return r;
    TRYC(event_sync(c.ev3), "svo_run");
    TRYC(hipEventElapsedTime(&ms_p, c.ev0, c.ev1), "svo_run");
    TRYC(hipEventElapsedTime(&ms_v, c.ev1, c.ev2), "svo_run");
    TRYC(hipEventElapsedTime(&ms_b, c.ev2, c.ev3), "svo_run");
  }
  if (info) {
    info->n_partitions = np;
    info->nfilled = c.nfilled;
    info->n_nodes = nn;
    info->n_data = nd;
    info->hazard = hz;
    info->ms_partition = ms_p;
    info->ms_voxelize = ms_v;
    info->ms_build = ms_b;
    info->launches = (uint32_t)(t_launches - l0);
    info->host_syncs = (uint32_t)(t_syncs - s0);
    info->graph_launches = info->graph_replays = info->graph_kernels = 0;
  }
  return SVO_OK;
}

int svo_run(svo_ctx* ctx, const float* records, uint64_t n_triangles, int payload,
            int src_on_device, const float bbox[6], uint64_t gridsize, uint64_t memory_limit_mb,
            int color_mode, int generate_levels, svo_run_info* info) {
  return run_impl(ctx, records, n_triangles, payload, src_on_device, bbox, gridsize,
                  memory_limit_mb, color_mode, generate_levels, false, nullptr, nullptr, info);
}

int svo_run_stream(svo_ctx* ctx, const float* records, uint64_t n_triangles, int payload,
                   int src_on_device, const float bbox[6], uint64_t gridsize,
                   uint64_t memory_limit_mb, int color_mode, int generate_levels,
                   svo_sink_fn sink, void* user, svo_run_info* info) {
  return run_impl(ctx, records, n_triangles, payload, src_on_device, bbox, gridsize,
                  memory_limit_mb, color_mode, generate_levels, true, sink, user, info);
}

// -levels jobs of a batch (OctreeBuilder.cpp:82-99 writes a data record per internal node,
// which the shared level passes do not): each runs alone through the plain path on the
// context's helper context, and its records are appended to the batch's arrays after the
// shared jobs'.
static int batch_with_levels(svo_ctx* ctx, const svo_job* jobs, uint32_t n_jobs,
                             svo_job_result* results, svo_run_info* info) {
  Context& c = ctx->c;
  const uint64_t l0 = t_launches, s0 = t_syncs;
  std::vector<svo_job> shared;
  std::vector<uint32_t> shared_idx, lv_idx;
  for (uint32_t j = 0; j < n_jobs; j++) {
    if (jobs[j].generate_levels) {
      lv_idx.push_back(j);
    } else {
      shared.push_back(jobs[j]);
      shared_idx.push_back(j);
    }
  }
  svo_run_info tot{};
  if (!shared.empty()) {
    std::vector<svo_job_result> res(shared.size());
    const int r = svo_batch_run(ctx, shared.data(), (uint32_t)shared.size(), res.data(), &tot);
    if (r) return r;
    for (size_t i = 0; i < shared.size(); i++) results[shared_idx[i]] = res[i];
  } else {
    reset_plain(c);
    c.batch_state = true;
  }
  if (!c.aux) {
    const int r = svo_ctx_create(c.device, (svo_ctx**)&c.aux);
    if (r) return r;
  }
  svo_ctx* aux = (svo_ctx*)c.aux;
  aux->c.tune = c.tune;
  aux->c.dense_limit = c.dense_limit;
  for (uint32_t j : lv_idx) {
    const svo_job& q = jobs[j];
    svo_run_info in{};
    int r = svo_run(aux, q.records, q.n_triangles, q.payload, q.src_on_device, q.bbox, q.gridsize,
                    q.memory_limit_mb, q.color_mode, 1, &in);
    if (r) {
      g_err = "batch job " + std::to_string(j) + ": " + g_err;
      return r;
    }
    CHECK_CTX();  // back on the batch context's device, error string cleared
    TRYC(ensure_keep(c.nodes, (c.n_nodes + in.n_nodes) * 24, c.n_nodes * 24, c.stream),
         "svo_batch_run");
    TRYC(ensure_keep(c.records, (c.n_data + in.n_data) * 32, c.n_data * 32, c.stream),
         "svo_batch_run");
    TRYC(hipMemcpyAsync((char*)c.nodes.p + c.n_nodes * 24, svo_nodes_device(aux),
                        in.n_nodes * 24, hipMemcpyDeviceToDevice, c.stream),
         "svo_batch_run");
    TRYC(hipMemcpyAsync((char*)c.records.p + c.n_data * 32, svo_data_device(aux),
                        in.n_data * 32, hipMemcpyDeviceToDevice, c.stream),
         "svo_batch_run");
    TRYC(stream_sync(c.stream), "svo_batch_run");
    svo_job_result& o = results[j];
    o.n_partitions = in.n_partitions;
    o.nfilled = in.nfilled;
    o.n_nodes = in.n_nodes;
    o.n_data = in.n_data;
    o.node_offset = c.n_nodes;
    o.data_offset = c.n_data;
    o.hazard = in.hazard;
    c.n_nodes += in.n_nodes;
    c.n_data += in.n_data;
    tot.n_partitions += in.n_partitions;
    tot.nfilled += in.nfilled;
    tot.hazard |= in.hazard;
    tot.ms_partition += in.ms_partition;
    tot.ms_voxelize += in.ms_voxelize;
    tot.ms_build += in.ms_build;
  }
  c.data_in_records2 = false;
  tot.n_nodes = c.n_nodes;
  tot.n_data = c.n_data;
  tot.launches = (uint32_t)(t_launches - l0);
  tot.host_syncs = (uint32_t)(t_syncs - s0);
  if (info) *info = tot;
  return SVO_OK;
}

int svo_batch_run(svo_ctx* ctx, const svo_job* jobs, uint32_t n_jobs, svo_job_result* results,
                  svo_run_info* info) {
  CHECK_CTX();
  Context& c = ctx->c;
  if (!jobs || !results || n_jobs == 0 || n_jobs > kMaxJobs) {
    g_err = "invalid batch arguments";
    return SVO_ERR_INVALID;
  }
  for (uint32_t j = 0; j < n_jobs; j++)
    if (jobs[j].generate_levels) return batch_with_levels(ctx, jobs, n_jobs, results, info);
  const uint64_t l0 = t_launches, s0 = t_syncs;
  uint64_t host_bytes = 0;
  for (uint32_t j = 0; j < n_jobs; j++) {
    const svo_job& q = jobs[j];
    if (!is_pow2_u64(q.gridsize) || q.gridsize > (1u << 15) || q.memory_limit_mb <= 1 ||
        q.color_mode < 0 || q.color_mode > 3 || (q.n_triangles && !q.records)) {
      g_err = "invalid arguments for batch job " + std::to_string(j);
      return SVO_ERR_INVALID;
    }
    if (!q.src_on_device) host_bytes += (q.n_triangles * (q.payload ? 84 : 36) + 255) & ~255ull;
  }
  // meshes in host memory are uploaded into one buffer; device meshes are used in place
  reset_plain(c);
  c.batch_state = true;
  const uint64_t gl0 = t_graph_launches, gr0 = t_graph_replays, gk0 = t_graph_kernels;
  GraphScope graph_scope(c, c.graph_policy != 0);
  if (host_bytes) TRYC(ensure(c.batch_tris, host_bytes), "svo_batch_run");
  c.jobs.assign(n_jobs, Job{});
  uint64_t hoff = 0;
  for (uint32_t j = 0; j < n_jobs; j++) {
    const svo_job& q = jobs[j];
    Job& J = c.jobs[j];
    const uint64_t bytes = q.n_triangles * (q.payload ? 84 : 36);
    if (q.src_on_device || !bytes) {
      J.tris = q.records;
    } else {
      J.tris = (const float*)((char*)c.batch_tris.p + hoff);
      TRYC(hipMemcpyAsync((void*)J.tris, q.records, bytes, hipMemcpyHostToDevice, c.stream),
           "svo_batch_run");
      hoff += (bytes + 255) & ~255ull;
    }
    J.n_tris = q.n_triangles;
    J.payload = q.payload ? 1 : 0;
    J.color_mode = q.color_mode;
    J.gridsize = q.gridsize;
    // main.cpp:306-311: the voxelizer's unitlength comes from the .trip text bbox
    J.unit_vox = (text_roundtrip(q.bbox[3]) - text_roundtrip(q.bbox[0])) / (float)q.gridsize;
  }
  TRYC(fill_u64(c.counters.as<uint64_t>(), 0, kNumCounters, c.stream), "svo_batch_run");
  float ms_p = 0, ms_v = 0, ms_b = 0;
  TRYC(hipEventRecord(c.ev0, c.stream), "svo_batch_run");
  uint64_t total_parts = 0;
  for (uint32_t j = 0; j < n_jobs; j++) {
    const svo_job& q = jobs[j];
    const uint64_t np = svo_estimate_partitions(q.gridsize, q.memory_limit_mb);
    const float u_part = (q.bbox[3] - q.bbox[0]) / (float)q.gridsize;  // partitioner.cpp:45
    TRYC(partition_job(c, j, u_part, q.gridsize, np), "svo_batch_run");
    total_parts += np;
  }
  TRYC(hipEventRecord(c.ev1, c.stream), "svo_batch_run");
  std::vector<VoxUnit> units;
  for (uint32_t j = 0; j < n_jobs; j++) {
    const Job& J = c.jobs[j];
    const uint64_t G = J.gridsize;
    const uint64_t mp = (G * G * G) / J.n_parts;
    for (uint64_t i = 0; i < J.n_parts; i++) {
      if (J.part_cnt[i] == 0) continue;  // main.cpp:330
      VoxUnit u;
      u.job = j;
      u.tris = J.tris;
      u.list = (J.n_parts > 1) ? c.part_lists.as<uint32_t>() + J.part_off[i] : nullptr;
      u.n = J.part_cnt[i];
      u.stride = J.payload ? 21 : 9;
      u.mstart = i * mp;
      u.mend = (i + 1) * mp;
      u.unit = J.unit_vox;
      u.gridsize = G;
      units.push_back(u);
    }
  }
  // passes of consecutive units whose Morton spans fit the dense byte grid together (a
  // larger single unit is a pass of its own, on the touched-brick path)
  for (Job& J : c.jobs) {
    J.voff = ~0ull;
    J.nvox = 0;
  }
  uint64_t nf = 0;
  for (size_t u0 = 0; u0 < units.size();) {
    size_t u1 = u0;
    uint64_t span = 0;
    while (u1 < units.size()) {
      const uint64_t bstart = units[u1].mstart & ~(uint64_t)(kBrickVox - 1);
      // dense cells: a byte per voxel, 4 bytes for payload meshes
      const uint64_t ub = (((units[u1].mend - bstart + kBrickVox - 1) >> kBrickBits)
                           << kBrickBits) * (units[u1].stride == 21 ? 4 : 1);
      if (u1 > u0 && span + ub > c.dense_limit) break;
      span += ub;
      u1++;
    }
    uint64_t nf1 = 0;
    TRYC(voxelize_units(c, std::vector<VoxUnit>(units.begin() + u0, units.begin() + u1), true,
                        &nf1),
         "svo_batch_run");
    nf += nf1;
    u0 = u1;
  }
  {  // jobs without voxels: an empty range where they would start
    uint64_t prev_end = 0;
    for (Job& J : c.jobs) {
      if (J.voff == ~0ull) J.voff = prev_end;
      prev_end = J.voff + J.nvox;
    }
  }
  c.nfilled = nf;
  TRYC(hipEventRecord(c.ev2, c.stream), "svo_batch_run");
  c.build_done_ev = c.ev3;  // recorded by the build before its final read-back
  hipError_t e = build_jobs(c, n_jobs, true, 0);
  c.build_done_ev = nullptr;
  if (e != hipSuccess) return fail(e, "svo_batch_run");
  TRYC(event_sync(c.ev3), "svo_batch_run");
  TRYC(hipEventElapsedTime(&ms_p, c.ev0, c.ev1), "svo_batch_run");
  TRYC(hipEventElapsedTime(&ms_v, c.ev1, c.ev2), "svo_batch_run");
  TRYC(hipEventElapsedTime(&ms_b, c.ev2, c.ev3), "svo_batch_run");
  unsigned long long err = 0;
  TRYC(read_err(c, &err),
       "svo_batch_run");
  if (err) {
    g_err = "internal: " + std::to_string(err) + kErrSources;
    return SVO_ERR_DEVICE;
  }
  int any_hazard = 0;
  for (uint32_t j = 0; j < n_jobs; j++) {
    const Job& J = c.jobs[j];
    svo_job_result& r = results[j];
    r.n_partitions = J.n_parts;
    r.nfilled = J.nvox;
    r.n_nodes = J.n_nodes;
    r.n_data = J.n_data;
    r.node_offset = J.nbase;
    r.data_offset = J.dbase;
    r.hazard = J.hazard;
    any_hazard |= J.hazard;
  }
  if (info) {
    info->n_partitions = total_parts;
    info->nfilled = nf;
    info->n_nodes = c.n_nodes;
    info->n_data = c.n_data;
    info->hazard = any_hazard;
    info->ms_partition = ms_p;
    info->ms_voxelize = ms_v;
    info->ms_build = ms_b;
    info->launches = (uint32_t)(t_launches - l0);
    info->host_syncs = (uint32_t)(t_syncs - s0);
    info->graph_launches = (uint32_t)(t_graph_launches - gl0);
    info->graph_replays = (uint32_t)(t_graph_replays - gr0);
    info->graph_kernels = (uint32_t)(t_graph_kernels - gk0);
  }
  return SVO_OK;
}

int svo_ctx_autotune(svo_ctx* ctx, const svo_job* jobs, uint32_t n_jobs, int best_out[4]) {
  CHECK_CTX();
  Context& c = ctx->c;
  std::vector<svo_job_result> res(n_jobs);
  // device time of the voxelize stage (the only one the knobs touch), best of two runs
  auto measure = [&](float* ms) -> int {
    float best = 1e30f;
    for (int rep = 0; rep < 2; rep++) {
      svo_run_info info;
      const int r = svo_batch_run(ctx, jobs, n_jobs, res.data(), &info);
      if (r) return r;
      best = std::min(best, info.ms_voxelize);
    }
    *ms = best;
    return SVO_OK;
  };
  float cur = 0;
  int r = measure(&cur);  // also the warm-up (allocations)
  if (r) return r;
  r = measure(&cur);
  if (r) return r;
  // coordinate descent over the three knobs; every setting gives the same output
  static const int cand_cols[] = {64, 256, 1024, 4096};
  static const int cand_direct[] = {0, 8, 16, 32};
  static const int cand_tile[] = {768, 1536, 2048, 4096};
  static const int cand_large[] = {1024, 2048, 4096};
  int* knob[4] = {&c.tune.small_cols, &c.tune.direct_cols, &c.tune.tile_blocks,
                  &c.tune.large_blocks};
  const int* cands[4] = {cand_cols, cand_direct, cand_tile, cand_large};
  const int ncand[4] = {4, 4, 4, 3};
  for (int k = 0; k < 4; k++) {
    const int keep = *knob[k];
    int best = keep;
    float best_ms = cur;
    for (int q = 0; q < ncand[k]; q++) {
      if (cands[k][q] == keep) continue;
      *knob[k] = cands[k][q];
      float ms = 0;
      r = measure(&ms);
      if (r) {
        *knob[k] = keep;
        return r;
      }
      if (ms < best_ms * 0.98f) {
        best_ms = ms;
        best = cands[k][q];
      }
    }
    *knob[k] = best;
    cur = best_ms;
  }
  if (best_out) {
    best_out[0] = c.tune.small_cols;
    best_out[1] = c.tune.direct_cols;
    best_out[2] = c.tune.tile_blocks;
    best_out[3] = c.tune.large_blocks;
  }
  return SVO_OK;
}

int svo_ctx_get_tuning(svo_ctx* ctx, int out[4]) {
  CHECK_CTX();
  if (!out) return SVO_ERR_INVALID;
  out[0] = ctx->c.tune.small_cols;
  out[1] = ctx->c.tune.direct_cols;
  out[2] = ctx->c.tune.tile_blocks;
  out[3] = ctx->c.tune.large_blocks;
  return SVO_OK;
}

int svo_ctx_set_knobs(svo_ctx* ctx, const int knobs[4]) {
  CHECK_CTX();
  if (!knobs || knobs[0] <= 0 || knobs[1] < 0 || knobs[2] <= 0 || knobs[3] <= 0) {
    g_err = "invalid voxelizer knobs";
    return SVO_ERR_INVALID;
  }
  ctx->c.tune.small_cols = knobs[0];
  ctx->c.tune.direct_cols = knobs[1];
  ctx->c.tune.tile_blocks = knobs[2];
  ctx->c.tune.large_blocks = knobs[3];
  return SVO_OK;
}

}  // extern "C"
