// ORACLE — test infrastructure, not product code.
//
// C interface of the single-threaded CPU restatement of ooc_svo_builder v1.6.4's
// partition -> voxelize -> SVO-build path (oracle/svo_oracle.cpp). Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as the
// checker or as the timed CPU baseline. The product library never links it.
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Morton codes (libmorton morton3D_64_encode/decode, default LUT build). */
void oracle_morton_encode_batch(const uint32_t* x, const uint32_t* y, const uint32_t* z,
                                size_t n, uint64_t* out);
void oracle_morton_decode_batch(const uint64_t* m, size_t n, uint32_t* x, uint32_t* y,
                                uint32_t* z);

/* partitioner.cpp:12-28 without the console output. */
size_t oracle_estimate_partitions(size_t gridsize, size_t memory_limit_mb);

/* partitioner.cpp:43-77,101-149: per-partition triangle counts and, when lists is
 * non-NULL, the in-order triangle indices of each partition concatenated (lists must
 * hold sum(counts) entries). tris: n records of 9 (binary) or 21 (payload) floats. */
void oracle_partition(const float* tris, size_t n, int payload, const float bbox[6],
                      size_t gridsize, size_t n_partitions, uint64_t* counts,
                      uint64_t* lists);

/* voxelizer.cpp:138-307 on one partition. Returns a handle; the voxel list is in
 * emission order (unsorted), as the reference's data vector. */
void* oracle_voxelize(const float* tris, size_t n, int payload, uint64_t morton_start,
                      uint64_t morton_end, float unitlength, float sparseness_limit);
size_t oracle_vox_count(void* h);
int oracle_vox_used_data(void* h);
void oracle_vox_codes(void* h, uint64_t* out);          /* binary and payload */
void oracle_vox_records(void* h, void* out32);          /* payload: 32 B records */
void oracle_vox_free(void* h);

/* OctreeBuilder.cpp:4-168 fed with sorted codes (binary) or sorted 32 B records
 * (payload). Returns a handle holding the .octreenodes/.octreedata bytes. */
void* oracle_build(const uint64_t* codes, const void* records32, size_t n, size_t gridsize,
                   int payload, int levels);
size_t oracle_build_nnodes(void* h);
size_t oracle_build_ndata(void* h);
int oracle_build_hazard(void* h);
void oracle_build_copy(void* h, void* nodes_out, void* data_out);
void oracle_build_free(void* h);

/* main.cpp:281-399 in memory: partition, per-partition voxelize, sort, colour mode,
 * build, finalize. bbox is the .tri header bbox. color_mode: 0 model, 1 fixed,
 * 2 linear, 3 normal. Returns a build handle (see oracle_build_*), and writes the
 * number of partitions and the total filled voxel count. */
void* oracle_pipeline(const float* tris, size_t n, int payload, const float bbox[6],
                      size_t gridsize, size_t memory_limit_mb, float sparseness_limit,
                      int color_mode, int levels, size_t* n_partitions_out,
                      size_t* nfilled_out);

/* The drop-in CLI (main.cpp) on files; payload selects svo_builder vs
 * svo_builder_binary behaviour. Returns the process exit status. */
int oracle_cli_main(int argc, char** argv, int payload);

#ifdef __cplusplus
}
#endif
