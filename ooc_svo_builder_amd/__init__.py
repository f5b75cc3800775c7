"""MI355X-native partition -> voxelize -> SVO-build pipeline (drop-in for
ooc_svo_builder's svo_builder / svo_builder_binary).

The compute path is libsvo_amd.so (HIP kernels for gfx950 behind the C ABI in
include/svo_amd.h); this package holds its Python binding, the file formats and the
synthetic workload generators.
"""
from .formats import (DATA_DTYPE, NODE_DTYPE, Octree, read_octree, read_tri,  # noqa: F401
                      write_tri)
from .lib import (CLI_BINARY, CLI_PAYLOAD, COLOR_MODES, LIB_PATH, SIGNATURES,  # noqa: F401
                  SvoContext, SvoError, SvoServer, estimate_partitions, load_library)
