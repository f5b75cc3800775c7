/*
 * svo_amd.h — C ABI of the MI355X partition -> voxelize -> SVO-build library
 * (libsvo_amd.so). Plain pointers and sizes only; every call returns SVO_OK (0) or an
 * error code, with the message in svo_last_error(). No exceptions cross the ABI.
 *
 * The entry points replace the call boundaries of ooc_svo_builder v1.6.4 that main.cpp
 * drives (paths relative to the reference repository):
 *   estimate_partitions      src/svo_builder/partitioner.h:12, partitioner.cpp:12-28
 *   partition                src/svo_builder/partitioner.h:14, partitioner.cpp:101-149
 *   voxelize_schwarz_method  src/svo_builder/voxelizer.h:25,27, voxelizer.cpp:138-307
 *   OctreeBuilder            src/svo_builder/OctreeBuilder.h:33-36, OctreeBuilder.cpp:4-168
 *   TriReader                src/libs/libtri/include/TriReader.h:9-83
 * The reference streams triangles from .tripdata files and appends nodes to files one
 * record at a time; here the mesh and every intermediate stay in HBM and the octree is
 * produced as two device arrays laid out exactly like .octreenodes / .octreedata.
 *
 * Threading: a context is used by one host thread at a time (thread-compatible, not
 * thread-safe); separate contexts are independent.
 */
#ifndef SVO_AMD_H
#define SVO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVO_OK 0
#define SVO_ERR_INVALID 1     /* bad argument or call order */
#define SVO_ERR_DEVICE 2      /* HIP runtime error */
#define SVO_ERR_NOMEM 3       /* device allocation failed */
#define SVO_ERR_UNSUPPORTED 4 /* option not implemented on the GPU path */
#define SVO_ERR_IO 5          /* the output sink (or an output file) failed */

/* colour modes of main.cpp:23,154-179 (-c model|fixed|linear|normal) */
#define SVO_COLOR_MODEL 0
#define SVO_COLOR_FIXED 1
#define SVO_COLOR_LINEAR 2
#define SVO_COLOR_NORMAL 3

typedef struct svo_ctx svo_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char* svo_last_error(void);
/* Library version string. */
const char* svo_version(void);

/* Context on HIP device 'device' with its own stream (no reference counterpart: the
 * reference has no device state). */
int svo_ctx_create(int device, svo_ctx** out);
void svo_ctx_destroy(svo_ctx* ctx);
/* Use a caller-owned hipStream_t (pass NULL to return to the context's own stream). */
int svo_ctx_set_stream(svo_ctx* ctx, void* hip_stream);
/* The hipStream_t the context launches on. */
void* svo_ctx_stream(svo_ctx* ctx);

/* Voxelizer tuning (no reference counterpart): small_cols = the column count above which
 * a triangle goes to the column-parallel tier (<= 0 keeps the current value);
 * dense_limit_mb = largest partition voxelized in a dense byte grid, larger ones use
 * touched bricks only (UINT64_MAX keeps the current value; the default is half of the
 * free HBM at context creation, at most 128 GiB). Neither changes the output. */
int svo_ctx_set_tuning(svo_ctx* ctx, int small_cols, uint64_t dense_limit_mb);

/* hipGraph use of svo_batch_run on this context (no reference counterpart; on by
 * default): each kernel sequence between two host synchronizations of a batch runs as a
 * graph, captured once and launched again by later batches that make the same launches
 * (same jobs, buffers and knobs). A context that captures must not share its stream, and
 * no other thread should synchronize the device or free device memory while it does
 * (captures are relaxed-mode); turn graphs off for such uses. Output is the same. */
int svo_ctx_set_graphs(svo_ctx* ctx, int on);

/* partitioner.cpp:12-28: number of Morton-range partitions (a power of 8) such that the
 * gridsize^3 byte grid fits memory_limit_mb. Pure integer rule, no device work. */
uint64_t svo_estimate_partitions(uint64_t gridsize, uint64_t memory_limit_mb);

/* Replaces TriReader (TriReader.h:9-83): load n_triangles .tridata records (9 floats
 * per triangle when payload == 0, 21 when payload == 1) into HBM. 'records' is a host
 * pointer, or a device pointer when src_on_device != 0. Resets the voxel stream. */
int svo_mesh_load(svo_ctx* ctx, const float* records, uint64_t n_triangles, int payload,
                  int src_on_device);

/* TriReader (TriReader.h:9-83) for a whole .tridata file: streams it into HBM through two
 * pinned 32 MiB chunks, the read of one overlapping the copy of the other. A file with
 * fewer than n_triangles records fails with SVO_ERR_INVALID; got_out (optional) receives
 * the complete records found. Resets the voxel stream like svo_mesh_load. */
int svo_mesh_load_file(svo_ctx* ctx, const char* tridata_path, uint64_t n_triangles,
                       int payload, uint64_t* got_out);

/* partitioner.cpp:101-149 on the device: for each of n_partitions Morton-range
 * partitions, the in-order list of triangles whose bbox touches the partition's world
 * box (unitlength from the .tri bbox, partitioner.cpp:45). counts_out[n_partitions]
 * receives the per-partition triangle counts written to the .trip header. */
int svo_partition(svo_ctx* ctx, float unitlength, uint64_t gridsize, uint64_t n_partitions,
                  uint64_t* counts_out);
/* Triangle indices of one partition (counts_out[partition] entries), for callers that
 * still write .tripdata files. */
int svo_partition_indices(svo_ctx* ctx, uint64_t partition, uint32_t* host_out);

/* partitioner.cpp:80-149 with trip_tools.h:104-125, for callers that still want the
 * reference's partition files on disk (the reference writes them as temporaries and
 * deletes them at exit, partitioner.cpp:31-40): after svo_partition, write
 * <base>.trip (the header, bbox printed with 6 significant digits, geo_only from the
 * loaded mesh) and <base>_<i>.tripdata with the records of partition i's triangles in mesh
 * order, for every non-empty partition (one partition: the whole mesh). base is the
 * reference's <tri base><gridsize>_<n_partitions>; bbox and version come from the .tri
 * header. */
int svo_write_trip(svo_ctx* ctx, const char* base, uint64_t gridsize, const float bbox[6],
                   int version);

/* voxelizer.cpp:138-307 followed by main.cpp:354-385: voxelize one partition (its list
 * from svo_partition, or the whole mesh when no partitioning was done) over Morton
 * codes [morton_start, morton_end) and append its voxels, in Morton order, to the
 * context's voxel stream. The range must be an aligned Morton cube (a partition: its
 * length a power of 8 dividing morton_start), as main.cpp:329-333 passes. color_mode
 * applies to payload meshes. nfilled_out receives the voxel count of this partition
 * (main.cpp:348). */
int svo_voxelize(svo_ctx* ctx, uint64_t partition, uint64_t morton_start,
                 uint64_t morton_end, float unitlength, uint64_t gridsize, int color_mode,
                 uint64_t* nfilled_out);

/* OctreeBuilder::addVoxel (OctreeBuilder.h:35-36, OctreeBuilder.cpp:131-168) in bulk:
 * append n voxels to the voxel stream. codes must be strictly increasing and above every
 * code already in the stream, as the reference's unsigned gap arithmetic requires
 * (OctreeBuilder.cpp:133-135). records32 (payload meshes only, else NULL) holds the n
 * 32-byte VoxelData records; all pointers are host memory. */
int svo_add_voxels(svo_ctx* ctx, const uint64_t* codes, const void* records32, uint64_t n);

/* OctreeBuilder (OctreeBuilder.cpp:4-168): build the octree of the voxel stream.
 * n_nodes/n_data are the .octree header counts (octree_io.h:74-83). hazard_out is set
 * to 1 when the reference's output would be undefined for this voxel set (last voxel
 * at the second-to-last Morton code, OctreeBuilder.cpp:36-42, or an empty 1^3 grid);
 * the completed tree is produced then. generate_levels != 0 (-levels) also writes a data
 * record for every internal node, the average of its children's (OctreeBuilder.cpp:82-99),
 * with every data record in post-order (n_data grows by the internal node count). */
int svo_build(svo_ctx* ctx, uint64_t gridsize, int generate_levels, uint64_t* n_nodes_out,
              uint64_t* n_data_out, int* hazard_out);

/* Copy records [first, first+count) of the built octree to 'host_dst' (host memory, or
 * device memory: the copy direction follows the pointer): nodes are 24 B
 * (u64 data, u64 children_base, i8 offsets[8]), data are 32 B (u64 morton, f32 color[3],
 * f32 normal[3]) — the .octreenodes / .octreedata layouts (octree_io.h:49-66). */
int svo_copy_nodes(svo_ctx* ctx, void* host_dst, uint64_t first, uint64_t count);
int svo_copy_data(svo_ctx* ctx, void* host_dst, uint64_t first, uint64_t count);
/* Device pointers of the built octree (valid until the next call on the context). */
const void* svo_nodes_device(svo_ctx* ctx);
const void* svo_data_device(svo_ctx* ctx);

/* Voxel stream: number of voxels so far, and a copy of their sorted Morton codes. */
uint64_t svo_voxel_count(svo_ctx* ctx);
int svo_copy_codes(svo_ctx* ctx, uint64_t* host_dst, uint64_t first, uint64_t count);

/* main.cpp:281-399 in memory: estimate partitions, partition, voxelize every non-empty
 * partition (unitlength from the bbox after the .trip text round trip, main.cpp:306-311),
 * build. bbox = the .tri header bbox (min xyz, max xyz). Outputs as for svo_build.
 * records == NULL reruns the pipeline on the mesh already in HBM (n_triangles and
 * src_on_device are then ignored; payload must match the loaded mesh). */
typedef struct svo_run_info {
  uint64_t n_partitions;
  uint64_t nfilled;
  uint64_t n_nodes;
  uint64_t n_data;
  int hazard;
  float ms_partition, ms_voxelize, ms_build; /* device time per stage */
  uint32_t launches;       /* kernel launches the call issued outside graphs */
  uint32_t host_syncs;     /* host waits on the device the call made */
  uint32_t graph_launches; /* hipGraph launches (svo_batch_run with graphs on) */
  uint32_t graph_replays;  /* ... of them, graphs kept from an earlier call */
  uint32_t graph_kernels;  /* kernels run inside those graphs */
} svo_run_info;
int svo_run(svo_ctx* ctx, const float* records, uint64_t n_triangles, int payload,
            int src_on_device, const float bbox[6], uint64_t gridsize,
            uint64_t memory_limit_mb, int color_mode, int generate_levels,
            svo_run_info* info);

/* Streamed (out-of-core) build: OctreeBuilder (OctreeBuilder.h:33-36,
 * OctreeBuilder.cpp:4-168) fed one partition at a time, as main.cpp:329-387 does, writing
 * its records out as each partition completes instead of holding the whole octree.
 *
 * The sink receives the .octreenodes (kind SVO_SINK_NODES, 24-byte records) and
 * .octreedata (kind SVO_SINK_DATA, 32-byte records) byte streams in order, in chunks of
 * whole records in host memory valid for the call; it returns 0 to go on, anything else
 * stops the build with SVO_ERR_IO. A NULL sink collects the records in HBM instead (then
 * svo_copy_nodes / svo_copy_data / svo_nodes_device read them after svo_stream_end).
 *
 * svo_stream_begin: the OctreeBuilder constructor (the null record, and the white record
 * of geometry-only builds, go out at once); gridsize and n_partitions as for svo_voxelize
 * (the mesh loaded by svo_mesh_load* decides payload or geometry-only).
 * svo_stream_partition: the voxels added since the previous call (svo_voxelize of this
 * partition, or svo_add_voxels within its Morton range) form the subtree of partition
 * 'partition': it is built, its records except its root go out, and the voxel stream is
 * emptied. Partitions come in increasing order; a partition never passed is empty.
 * svo_stream_end: finalizeTree (OctreeBuilder.cpp:34-55): the partition roots, the
 * levels above them and the root go out; the .octree header counts and the hazard flag
 * (as for svo_build). No part of the build holds more than one partition's voxels and
 * subtree, so the 2^32-voxel limit of svo_build applies per partition. */
#define SVO_SINK_NODES 0
#define SVO_SINK_DATA 1
typedef int (*svo_sink_fn)(void* user, int kind, const void* bytes, uint64_t size);
int svo_stream_begin(svo_ctx* ctx, uint64_t gridsize, uint64_t n_partitions,
                     int generate_levels, svo_sink_fn sink, void* user);
int svo_stream_partition(svo_ctx* ctx, uint64_t partition);
int svo_stream_end(svo_ctx* ctx, uint64_t* n_nodes_out, uint64_t* n_data_out, int* hazard_out);

/* main.cpp:281-399 with the streamed build: like svo_run, the records going to 'sink'
 * (see svo_stream_begin) partition by partition. svo_run itself takes this path, with the
 * records collected in HBM, whenever the grid is split into more than one partition. */
int svo_run_stream(svo_ctx* ctx, const float* records, uint64_t n_triangles, int payload,
                   int src_on_device, const float bbox[6], uint64_t gridsize,
                   uint64_t memory_limit_mb, int color_mode, int generate_levels,
                   svo_sink_fn sink, void* user, svo_run_info* info);

/* Batched form (no reference counterpart: main.cpp builds one mesh per process).
 * Runs main.cpp:281-399 for n_jobs meshes at once: every job's partitions are voxelized
 * in the same persistent launches (one device work queue of 256-triangle tiles over all
 * meshes) and all octrees are built by the same level passes, so the launch count does
 * not grow with the number of meshes. Each job's output is byte-identical to svo_run on
 * that job alone. Results: the node and data records of job j are records
 * [node_offset, node_offset + n_nodes) and [data_offset, data_offset + n_data) of the
 * batch's arrays (svo_copy_nodes / svo_copy_data / svo_nodes_device / svo_data_device).
 * Limits: n_jobs <= 65536, gridsize <= 32768, fewer than 2^32 triangle references and
 * voxels in the whole batch. Jobs with generate_levels (-levels, whose internal nodes
 * carry averaged data records) run one at a time through the plain path after the
 * others, their records appended to the batch's arrays. 'records' of a job is a host
 * pointer, or a device pointer when src_on_device != 0 (used in place, no copy). */
typedef struct svo_job {
  const float* records;
  uint64_t n_triangles;
  int payload;
  int src_on_device;
  float bbox[6]; /* .tri header bbox: min xyz, max xyz */
  uint64_t gridsize;
  uint64_t memory_limit_mb;
  int color_mode;
  int generate_levels;
} svo_job;
typedef struct svo_job_result {
  uint64_t n_partitions;
  uint64_t nfilled;
  uint64_t n_nodes;
  uint64_t n_data;
  uint64_t node_offset;
  uint64_t data_offset;
  int hazard;
} svo_job_result;
/* info (optional) receives batch totals and the device time of each stage, and the
 * launch counts (see svo_ctx_set_graphs for the graphs). */
int svo_batch_run(svo_ctx* ctx, const svo_job* jobs, uint32_t n_jobs, svo_job_result* results,
                  svo_run_info* info);

/* Autotuner (no reference counterpart): runs the batch once per candidate value of the
 * voxelizer's knobs — the small/large-triangle split threshold, the column count up to
 * which a triangle's own thread walks it, and the grid sizes of the persistent
 * small-tier launch and of the large tier — by coordinate descent, keeping the setting
 * with the lowest voxelize device time on the context. No knob changes any output.
 * best_out (optional) receives {small_cols, direct_cols, tile_blocks, large_blocks};
 * the context's batch outputs afterwards are those of the last run. */
int svo_ctx_autotune(svo_ctx* ctx, const svo_job* jobs, uint32_t n_jobs, int best_out[4]);
/* The context's current knobs {small_cols, direct_cols, tile_blocks, large_blocks}, and
 * setting them explicitly. */
int svo_ctx_get_tuning(svo_ctx* ctx, int out[4]);
int svo_ctx_set_knobs(svo_ctx* ctx, const int knobs[4]);

/* Served form (no reference counterpart). A server owns 'slots' contexts, each with its
 * own stream and a worker thread. Submitted jobs queue in arrival order; a free slot
 * takes, from the head of the queue, as many jobs as fit the admission budget (estimated
 * HBM footprint of each job's working set and of its held results) and max_batch_jobs,
 * and runs them as one batch (svo_batch_run). Jobs that arrive while batches are on the
 * GPU are admitted by the next free slot, so batches of different slots overlap on the
 * device. A job that finds no other job queued, and every -levels job, runs alone
 * through the plain path (svo_run). A job's node and data records stay in HBM, owned by
 * its ticket, until svo_server_release. Every result is byte-identical to svo_run on that
 * job alone.
 * All svo_server_* calls are thread-safe. A host 'records' pointer must stay valid until
 * the job is done (the records are read when its batch starts). */
typedef struct svo_server svo_server;
/* admission_budget_mb = 0 selects 200 GiB (at most three quarters of the device memory
 * free at creation); max_batch_jobs = 0 selects 512; slots = 0 selects 2. */
int svo_server_create(int device, uint64_t admission_budget_mb, uint32_t max_batch_jobs,
                      uint32_t slots, svo_server** out);
/* Waits for running batches, then frees every ticket and context. */
void svo_server_destroy(svo_server* srv);
/* Queue a job; returns at once with its ticket (> 0). */
int svo_server_submit(svo_server* srv, const svo_job* job, uint64_t* ticket_out);
/* state_out: 0 queued, 1 running, 2 done, 3 failed. */
#define SVO_JOB_QUEUED 0
#define SVO_JOB_RUNNING 1
#define SVO_JOB_DONE 2
#define SVO_JOB_FAILED 3
int svo_server_poll(svo_server* srv, uint64_t ticket, int* state_out);
/* Block until the job is done or failed. Returns the job's own status (SVO_OK or the
 * batch's error code, message in svo_last_error()); result_out (optional) receives its
 * counts (node_offset = data_offset = 0: the records are the ticket's own arrays) and
 * latency_ms (optional) the time from submission to completion. */
int svo_server_wait(svo_server* srv, uint64_t ticket, svo_job_result* result_out,
                    double* latency_ms);
/* The ticket's node / data records: copies (dst in host or device memory) and device
 * pointers (valid until svo_server_release). */
int svo_server_copy_nodes(svo_server* srv, uint64_t ticket, void* dst, uint64_t first,
                          uint64_t count);
int svo_server_copy_data(svo_server* srv, uint64_t ticket, void* dst, uint64_t first,
                         uint64_t count);
const void* svo_server_nodes_device(svo_server* srv, uint64_t ticket);
const void* svo_server_data_device(svo_server* srv, uint64_t ticket);
/* Free a finished ticket and its records (returns its HBM to the admission budget). */
int svo_server_release(svo_server* srv, uint64_t ticket);
typedef struct svo_server_stats {
  uint64_t submitted, completed, failed, batches;
  uint64_t max_batch_jobs_seen;
  uint64_t peak_admitted_bytes;
  uint64_t batched_jobs; /* jobs over all batches */
  double batch_ms;       /* wall time the slots spent running batches (summed) */
  uint64_t fallback_batches; /* batches that failed and ran job by job */
  uint64_t buffer_retries;   /* result-buffer allocations that emptied the cache first */
  uint64_t admitted_bytes;   /* now: running estimates + held results */
  uint64_t held_bytes;       /* now: results held until released */
  uint64_t queued;           /* now: jobs waiting for admission */
  double shared_ms;  /* wall time in the shared batches (svo_batch_run), summed */
  double device_ms;  /* ... of which their stages' device time */
  double collect_ms; /* wall time copying results into the tickets' buffers */
  double alone_ms;   /* wall time of jobs run alone (svo_run) */
  uint64_t budget_bytes; /* the admission budget in force (the request, capped) */
} svo_server_stats;
int svo_server_stats_get(svo_server* srv, svo_server_stats* out);

/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc). The device
 * work of svo_partition (partitioner.cpp:43-77,101-149) on a mesh already in device
 * memory, without a context: its working memory is a fixed array in the library's device
 * image (about 9 MiB per device, held until the process ends), so the call allocates
 * nothing and does not synchronize; it is ordered on hip_stream (NULL: the null stream)
 * of the calling thread's current device.
 * d_records: n_triangles <= SVO_NOALLOC_MAX_TRIANGLES device records (9 or 21 floats);
 * n_partitions a power of 8 up to SVO_NOALLOC_MAX_PARTITIONS. Results, in device memory:
 * d_lists[lists_capacity] the per-partition triangle lists concatenated in partition
 * order (partition j's list starts at the sum of the counts before j; one partition lists
 * every triangle, as svo_partition does), d_counts[n_partitions + 1] the counts and then a
 * status word: 0 when every list is complete, 1 when they exceed lists_capacity (entries
 * beyond it are dropped), 2 on an internal scan failure. Calls that overlap on one device
 * share its working memory and race: run them one after another. */
#define SVO_NOALLOC_MAX_TRIANGLES (1u << 20)
#define SVO_NOALLOC_MAX_PARTITIONS 4096u
int svo_partition_noalloc(void* hip_stream, const float* d_records, uint64_t n_triangles,
                          int payload, float unitlength, uint64_t gridsize,
                          uint64_t n_partitions, uint32_t* d_lists, uint64_t lists_capacity,
                          uint64_t* d_counts);

/* Per-kernel device timing with HIP events on the context's stream, the counterpart of
 * the reference's stage timers (globals.h:8-29, main.cpp:45-56; those truncate to whole
 * milliseconds, timer.h:59). While on, each instrumented kernel launch is bracketed by
 * two events; svo_kernel_timing returns the launch count and summed milliseconds for one
 * kernel name (e.g. "vox_small_mark"). */
int svo_ctx_set_kernel_timing(svo_ctx* ctx, int on);
int svo_kernel_timing(svo_ctx* ctx, const char* name, uint64_t* launches, double* total_ms);
int svo_kernel_timing_reset(svo_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* SVO_AMD_H */
