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

/* partitioner.cpp:12-28: number of Morton-range partitions (a power of 8) such that the
 * gridsize^3 byte grid fits memory_limit_mb. Pure integer rule, no device work. */
uint64_t svo_estimate_partitions(uint64_t gridsize, uint64_t memory_limit_mb);

/* Replaces TriReader (TriReader.h:9-83): load n_triangles .tridata records (9 floats
 * per triangle when payload == 0, 21 when payload == 1) into HBM. 'records' is a host
 * pointer, or a device pointer when src_on_device != 0. Resets the voxel stream. */
int svo_mesh_load(svo_ctx* ctx, const float* records, uint64_t n_triangles, int payload,
                  int src_on_device);

/* partitioner.cpp:101-149 on the device: for each of n_partitions Morton-range
 * partitions, the in-order list of triangles whose bbox touches the partition's world
 * box (unitlength from the .tri bbox, partitioner.cpp:45). counts_out[n_partitions]
 * receives the per-partition triangle counts written to the .trip header. */
int svo_partition(svo_ctx* ctx, float unitlength, uint64_t gridsize, uint64_t n_partitions,
                  uint64_t* counts_out);
/* Triangle indices of one partition (counts_out[partition] entries), for callers that
 * still write .tripdata files. */
int svo_partition_indices(svo_ctx* ctx, uint64_t partition, uint32_t* host_out);

/* voxelizer.cpp:138-307 followed by main.cpp:354-385: voxelize one partition (its list
 * from svo_partition, or the whole mesh when no partitioning was done) over Morton
 * codes [morton_start, morton_end) and append its voxels, in Morton order, to the
 * context's voxel stream. color_mode applies to payload meshes. nfilled_out receives
 * the voxel count of this partition (main.cpp:348). */
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
 * the completed tree is produced then. generate_levels (-levels) is not supported yet
 * and returns SVO_ERR_UNSUPPORTED. */
int svo_build(svo_ctx* ctx, uint64_t gridsize, int generate_levels, uint64_t* n_nodes_out,
              uint64_t* n_data_out, int* hazard_out);

/* Copy records [first, first+count) of the built octree to host memory: nodes are 24 B
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
} svo_run_info;
int svo_run(svo_ctx* ctx, const float* records, uint64_t n_triangles, int payload,
            int src_on_device, const float bbox[6], uint64_t gridsize,
            uint64_t memory_limit_mb, int color_mode, int generate_levels,
            svo_run_info* info);

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
