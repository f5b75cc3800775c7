"""ctypes binding of libsvo_amd.so (C ABI in include/svo_amd.h).

The library is built in-tree (ooc_svo_builder_amd/lib/libsvo_amd.so) by
``__graft_entry__.build()``; loading fails loudly when it is missing — there is no
CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .formats import DATA_DTYPE, NODE_DTYPE

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libsvo_amd.so")
CLI_PAYLOAD = os.path.join(PKG, "bin", "svo_builder")
CLI_BINARY = os.path.join(PKG, "bin", "svo_builder_binary")
OCTREE_TOOL = os.path.join(PKG, "bin", "svo_octree_tool")
TRI_CONVERT = os.path.join(PKG, "bin", "tri_convert")
TRI_CONVERT_BINARY = os.path.join(PKG, "bin", "tri_convert_binary")

SVO_OK = 0
SVO_ERR_INVALID = 1
SVO_ERR_DEVICE = 2
SVO_ERR_NOMEM = 3
SVO_ERR_UNSUPPORTED = 4
SVO_ERR_IO = 5

SINK_NODES = 0
SINK_DATA = 1
# svo_sink_fn: int (*)(void* user, int kind, const void* bytes, uint64_t size)
SINK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_uint64)

COLOR_MODES = {"model": 0, "fixed": 1, "linear": 2, "normal": 3}

# Every function declared in include/svo_amd.h, with (restype, argtypes).
_vp, _u64, _i32, _f32 = C.c_void_p, C.c_uint64, C.c_int, C.c_float


class RunInfo(C.Structure):
    _fields_ = [("n_partitions", C.c_uint64), ("nfilled", C.c_uint64),
                ("n_nodes", C.c_uint64), ("n_data", C.c_uint64), ("hazard", C.c_int),
                ("ms_partition", C.c_float), ("ms_voxelize", C.c_float),
                ("ms_build", C.c_float), ("launches", C.c_uint32), ("host_syncs", C.c_uint32),
                ("graph_launches", C.c_uint32), ("graph_replays", C.c_uint32),
                ("graph_kernels", C.c_uint32)]


class Job(C.Structure):
    """svo_job: one mesh of a batch."""
    _fields_ = [("records", C.c_void_p), ("n_triangles", C.c_uint64), ("payload", C.c_int),
                ("src_on_device", C.c_int), ("bbox", C.c_float * 6), ("gridsize", C.c_uint64),
                ("memory_limit_mb", C.c_uint64), ("color_mode", C.c_int),
                ("generate_levels", C.c_int)]


class JobResult(C.Structure):
    """svo_job_result: a batch job's counts and its ranges in the batch's arrays."""
    _fields_ = [("n_partitions", C.c_uint64), ("nfilled", C.c_uint64),
                ("n_nodes", C.c_uint64), ("n_data", C.c_uint64),
                ("node_offset", C.c_uint64), ("data_offset", C.c_uint64), ("hazard", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("batches", C.c_uint64), ("max_batch_jobs_seen", C.c_uint64),
                ("peak_admitted_bytes", C.c_uint64), ("batched_jobs", C.c_uint64),
                ("batch_ms", C.c_double), ("fallback_batches", C.c_uint64),
                ("buffer_retries", C.c_uint64), ("admitted_bytes", C.c_uint64),
                ("held_bytes", C.c_uint64), ("queued", C.c_uint64),
                ("shared_ms", C.c_double), ("device_ms", C.c_double),
                ("collect_ms", C.c_double), ("alone_ms", C.c_double),
                ("budget_bytes", C.c_uint64)]


SIGNATURES = {
    "svo_last_error": (C.c_char_p, []),
    "svo_version": (C.c_char_p, []),
    "svo_ctx_create": (_i32, [_i32, C.POINTER(_vp)]),
    "svo_ctx_destroy": (None, [_vp]),
    "svo_ctx_set_stream": (_i32, [_vp, _vp]),
    "svo_ctx_stream": (_vp, [_vp]),
    "svo_ctx_set_tuning": (_i32, [_vp, _i32, _u64]),
    "svo_ctx_set_graphs": (_i32, [_vp, _i32]),
    "svo_estimate_partitions": (_u64, [_u64, _u64]),
    "svo_mesh_load": (_i32, [_vp, _vp, _u64, _i32, _i32]),
    "svo_mesh_load_file": (_i32, [_vp, C.c_char_p, _u64, _i32, C.POINTER(_u64)]),
    "svo_partition": (_i32, [_vp, _f32, _u64, _u64, _vp]),
    "svo_partition_indices": (_i32, [_vp, _u64, _vp]),
    "svo_write_trip": (_i32, [_vp, C.c_char_p, _u64, _vp, _i32]),
    "svo_voxelize": (_i32, [_vp, _u64, _u64, _u64, _f32, _u64, _i32, C.POINTER(_u64)]),
    "svo_add_voxels": (_i32, [_vp, _vp, _vp, _u64]),
    "svo_build": (_i32, [_vp, _u64, _i32, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_i32)]),
    "svo_copy_nodes": (_i32, [_vp, _vp, _u64, _u64]),
    "svo_copy_data": (_i32, [_vp, _vp, _u64, _u64]),
    "svo_nodes_device": (_vp, [_vp]),
    "svo_data_device": (_vp, [_vp]),
    "svo_voxel_count": (_u64, [_vp]),
    "svo_copy_codes": (_i32, [_vp, _vp, _u64, _u64]),
    "svo_ctx_set_kernel_timing": (_i32, [_vp, _i32]),
    "svo_kernel_timing": (_i32, [_vp, C.c_char_p, C.POINTER(_u64), C.POINTER(C.c_double)]),
    "svo_kernel_timing_reset": (_i32, [_vp]),
    "svo_run": (_i32, [_vp, _vp, _u64, _i32, _i32, _vp, _u64, _u64, _i32, _i32,
                       C.POINTER(RunInfo)]),
    "svo_stream_begin": (_i32, [_vp, _u64, _u64, _i32, _vp, _vp]),
    "svo_stream_partition": (_i32, [_vp, _u64]),
    "svo_stream_end": (_i32, [_vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_i32)]),
    "svo_run_stream": (_i32, [_vp, _vp, _u64, _i32, _i32, _vp, _u64, _u64, _i32, _i32, _vp,
                              _vp, C.POINTER(RunInfo)]),
    "svo_partition_noalloc": (_i32, [_vp, _vp, _u64, _i32, _f32, _u64, _u64, _vp, _u64, _vp]),
    "svo_batch_run": (_i32, [_vp, C.POINTER(Job), C.c_uint32, C.POINTER(JobResult),
                             C.POINTER(RunInfo)]),
    "svo_ctx_autotune": (_i32, [_vp, C.POINTER(Job), C.c_uint32, C.POINTER(C.c_int * 4)]),
    "svo_ctx_get_tuning": (_i32, [_vp, C.POINTER(C.c_int * 4)]),
    "svo_ctx_set_knobs": (_i32, [_vp, C.POINTER(C.c_int * 4)]),
    "svo_server_create": (_i32, [_i32, _u64, C.c_uint32, C.c_uint32, C.POINTER(_vp)]),
    "svo_server_destroy": (None, [_vp]),
    "svo_server_submit": (_i32, [_vp, C.POINTER(Job), C.POINTER(_u64)]),
    "svo_server_poll": (_i32, [_vp, _u64, C.POINTER(_i32)]),
    "svo_server_wait": (_i32, [_vp, _u64, C.POINTER(JobResult), C.POINTER(C.c_double)]),
    "svo_server_copy_nodes": (_i32, [_vp, _u64, _vp, _u64, _u64]),
    "svo_server_copy_data": (_i32, [_vp, _u64, _vp, _u64, _u64]),
    "svo_server_nodes_device": (_vp, [_vp, _u64]),
    "svo_server_data_device": (_vp, [_vp, _u64]),
    "svo_server_release": (_i32, [_vp, _u64]),
    "svo_server_stats_get": (_i32, [_vp, C.POINTER(ServerStats)]),
}

_lib = None


class SvoError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("svo error %d: %s" % (code, msg))
        self.code = code


def load_library(path: str = LIB_PATH):
    """Load libsvo_amd.so and declare its signatures; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libsvo_amd.so is not built (%s); run __graft_entry__.build()" % path)
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc != SVO_OK:
        raise SvoError(rc, load_library().svo_last_error().decode())


def estimate_partitions(gridsize: int, memory_limit_mb: int) -> int:
    return int(load_library().svo_estimate_partitions(gridsize, memory_limit_mb))


def partition_noalloc(stream_ptr: int, records, payload: bool, unitlength: float, gridsize: int,
                      n_partitions: int, lists, counts):
    """svo_partition_noalloc on device tensors: records (N, 9|21) float32, lists int32/uint32
    (its length is the capacity), counts int64 of n_partitions + 1 entries. Asynchronous
    on the given stream."""
    ptr, n, _, _keep = _mesh(records, payload)
    _check(load_library().svo_partition_noalloc(
        C.c_void_p(stream_ptr), ptr, n,
        int(payload), float(np.float32(unitlength)), gridsize, n_partitions,
        C.c_void_p(lists.data_ptr()), int(lists.numel()), C.c_void_p(counts.data_ptr())))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _mesh(records, payload: bool):
    """(pointer, triangle count, on device, object to keep alive) of a mesh given as an
    (N, 9) binary or (N, 21) payload float32 array: a numpy array or a contiguous device
    tensor. The C ABI reads N records of that width, so any other shape is refused here."""
    width = 21 if payload else 9
    if hasattr(records, "data_ptr"):
        if records.dim() != 2 or records.shape[1] != width or \
                str(records.dtype) != "torch.float32" or not records.is_contiguous():
            raise ValueError("mesh records must be a contiguous (N, %d) float32 tensor for "
                             "payload=%s, got %s %s" % (width, bool(payload),
                                                        tuple(records.shape), records.dtype))
        return (C.c_void_p(records.data_ptr()), int(records.shape[0]),
                1 if records.is_cuda else 0, records)
    r = np.ascontiguousarray(records, np.float32)
    if r.ndim != 2 or r.shape[1] != width:
        raise ValueError("mesh records must be an (N, %d) float32 array for payload=%s, got "
                         "shape %s" % (width, bool(payload), r.shape))
    return _ptr(r), int(r.shape[0]), 0, r


def fill_job(q: Job, j: dict):
    """Fill an svo_job from a dict with keys records ((N, 9|21) float32 numpy array, or a
    device tensor), payload, gridsize and optionally memory_limit_mb (2048), bbox
    ((0,0,0,1,1,1)), color_mode (0) and levels (False). Returns the object that must stay
    alive while the job's records are read."""
    ptr, n, on_dev, keep = _mesh(j["records"], j["payload"])
    q.records = ptr.value
    q.n_triangles = n
    q.src_on_device = on_dev
    q.payload = int(bool(j["payload"]))
    for k, v in enumerate(j.get("bbox", (0, 0, 0, 1, 1, 1))):
        q.bbox[k] = float(v)
    q.gridsize = int(j["gridsize"])
    q.memory_limit_mb = int(j.get("memory_limit_mb", 2048))
    q.color_mode = int(j.get("color_mode", 0))
    q.generate_levels = int(bool(j.get("levels", False)))
    return keep


class SvoServer:
    """Served form: submit jobs as they arrive; results stay in HBM per ticket."""

    def __init__(self, device: int = 0, budget_mb: int = 0, max_batch: int = 0, slots: int = 0):
        L = load_library()
        h = C.c_void_p()
        _check(L.svo_server_create(device, budget_mb, max_batch, slots, C.byref(h)))
        self._h = h
        self._keep = {}

    def close(self):
        if self._h:
            load_library().svo_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, job: dict) -> int:
        q = Job()
        keep = fill_job(q, job)
        t = C.c_uint64(0)
        _check(load_library().svo_server_submit(self._h, C.byref(q), C.byref(t)))
        self._keep[t.value] = keep
        return int(t.value)

    def poll(self, ticket: int) -> int:
        st = C.c_int(0)
        _check(load_library().svo_server_poll(self._h, ticket, C.byref(st)))
        return int(st.value)

    def wait(self, ticket: int):
        """(JobResult, latency in ms) once the job is done; raises if it failed."""
        r = JobResult()
        lat = C.c_double(0)
        rc = load_library().svo_server_wait(self._h, ticket, C.byref(r), C.byref(lat))
        self._keep.pop(ticket, None)
        _check(rc)
        return r, float(lat.value)

    def nodes(self, ticket: int, n: int) -> np.ndarray:
        out = np.empty(n, NODE_DTYPE)
        _check(load_library().svo_server_copy_nodes(self._h, ticket, _ptr(out), 0, n))
        return out

    def data(self, ticket: int, n: int) -> np.ndarray:
        out = np.empty(n, DATA_DTYPE)
        _check(load_library().svo_server_copy_data(self._h, ticket, _ptr(out), 0, n))
        return out

    def copy_nodes_to(self, ticket: int, dst_ptr: int, n: int):
        _check(load_library().svo_server_copy_nodes(self._h, ticket, C.c_void_p(dst_ptr), 0, n))

    def copy_data_to(self, ticket: int, dst_ptr: int, n: int):
        _check(load_library().svo_server_copy_data(self._h, ticket, C.c_void_p(dst_ptr), 0, n))

    def release(self, ticket: int):
        _check(load_library().svo_server_release(self._h, ticket))

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(load_library().svo_server_stats_get(self._h, C.byref(st)))
        return st


class SvoContext:
    """One device context: mesh, partitions, voxel stream and built octree in HBM."""

    def __init__(self, device: int = 0):
        L = load_library()
        h = C.c_void_p()
        _check(L.svo_ctx_create(device, C.byref(h)))
        self._h = h
        self.payload = False

    def close(self):
        if self._h:
            load_library().svo_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return int(load_library().svo_ctx_stream(self._h) or 0)

    def set_stream(self, stream_ptr: int | None):
        _check(load_library().svo_ctx_set_stream(self._h, C.c_void_p(stream_ptr or 0)))

    def set_tuning(self, small_cols: int = 0, dense_limit_mb: int | None = None):
        """Voxelizer split threshold and dense-grid limit (output is unchanged)."""
        lim = 0xFFFFFFFFFFFFFFFF if dense_limit_mb is None else int(dense_limit_mb)
        _check(load_library().svo_ctx_set_tuning(self._h, int(small_cols), lim))

    def set_graphs(self, on: bool):
        """hipGraph capture and replay of batch_run's kernel sequences (on by default)."""
        _check(load_library().svo_ctx_set_graphs(self._h, 1 if on else 0))

    # -- stages -----------------------------------------------------------------
    def load_mesh(self, records, payload: bool):
        """records: (N, 9|21) float32 numpy array, or a torch tensor on the device."""
        ptr, n, on_dev, _keep = _mesh(records, payload)
        _check(load_library().svo_mesh_load(self._h, ptr, n, int(payload), on_dev))
        self.payload = bool(payload)

    def load_mesh_file(self, tridata_path: str, n_triangles: int, payload: bool) -> int:
        """Stream a .tridata file into HBM; returns the triangle count loaded."""
        got = C.c_uint64(0)
        _check(load_library().svo_mesh_load_file(self._h, tridata_path.encode(), n_triangles,
                                                 int(payload), C.byref(got)))
        self.payload = bool(payload)
        return int(got.value)

    def partition(self, unitlength: float, gridsize: int, n_partitions: int) -> np.ndarray:
        counts = np.zeros(n_partitions, np.uint64)
        _check(load_library().svo_partition(self._h, float(np.float32(unitlength)), gridsize,
                                            n_partitions, _ptr(counts)))
        return counts

    def write_trip(self, base: str, gridsize: int, bbox=(0, 0, 0, 1, 1, 1), version: int = 1):
        """Write <base>.trip and <base>_<i>.tripdata after partition()."""
        b = np.asarray(bbox, np.float32)
        _check(load_library().svo_write_trip(self._h, base.encode(), gridsize, _ptr(b), version))

    def partition_indices(self, part: int, count: int) -> np.ndarray:
        out = np.empty(count, np.uint32)
        _check(load_library().svo_partition_indices(self._h, part, _ptr(out)))
        return out

    def voxelize(self, part: int, morton_start: int, morton_end: int, unitlength: float,
                 gridsize: int, color_mode: int = 0) -> int:
        nf = C.c_uint64(0)
        _check(load_library().svo_voxelize(self._h, part, morton_start, morton_end,
                                           float(np.float32(unitlength)), gridsize, color_mode,
                                           C.byref(nf)))
        return int(nf.value)

    def add_voxels(self, codes: np.ndarray, records: np.ndarray | None = None):
        c = np.ascontiguousarray(codes, np.uint64)
        r = None if records is None else np.ascontiguousarray(records)
        _check(load_library().svo_add_voxels(self._h, _ptr(c), None if r is None else _ptr(r),
                                             c.size))

    def build(self, gridsize: int, levels: bool = False):
        nn, nd, hz = C.c_uint64(0), C.c_uint64(0), C.c_int(0)
        _check(load_library().svo_build(self._h, gridsize, int(levels), C.byref(nn),
                                        C.byref(nd), C.byref(hz)))
        return int(nn.value), int(nd.value), int(hz.value)

    def run(self, records, payload: bool, gridsize: int, memory_limit_mb: int = 2048,
            bbox=(0, 0, 0, 1, 1, 1), color_mode: int = 0, levels: bool = False) -> RunInfo:
        """main.cpp's flow in memory; the mesh may be a numpy array or a device tensor."""
        b = np.asarray(bbox, np.float32)
        info = RunInfo()
        ptr, n, on_dev, _keep = _mesh(records, payload)
        _check(load_library().svo_run(self._h, ptr, n, int(payload), on_dev, _ptr(b), gridsize,
                                      memory_limit_mb, color_mode, int(levels), C.byref(info)))
        self.payload = bool(payload)
        self.last_info = info
        return info

    def run_stream(self, records, payload: bool, gridsize: int, sink, memory_limit_mb: int = 2048,
                   bbox=(0, 0, 0, 1, 1, 1), color_mode: int = 0, levels: bool = False) -> RunInfo:
        """main.cpp's flow with the streamed build: sink(kind, memoryview) receives the
        .octreenodes (kind 0) and .octreedata (kind 1) bytes in order and returns nothing
        (an exception stops the build). records=None reruns on the mesh in HBM."""
        L = load_library()
        b = np.asarray(bbox, np.float32)
        info = RunInfo()
        errors = []

        def _cb(user, kind, ptr, size):
            try:
                sink(kind, (C.c_char * size).from_address(ptr))
                return 0
            except Exception as e:  # reported after the call
                errors.append(e)
                return 1
        cb = SINK_FN(_cb)
        _keep = None
        if records is None:
            args = (None, 0, int(payload), 0)
        else:
            ptr, n, on_dev, _keep = _mesh(records, payload)
            args = (ptr, n, int(payload), on_dev)
        rc = L.svo_run_stream(self._h, *args, _ptr(b), gridsize, memory_limit_mb, color_mode,
                              int(levels), C.cast(cb, C.c_void_p), None, C.byref(info))
        if errors:
            raise errors[0]
        _check(rc)
        self.payload = bool(payload)
        self.last_info = info
        return info

    def stream_begin(self, gridsize: int, n_partitions: int, levels: bool = False):
        """Streamed build with its records collected in HBM (no sink)."""
        _check(load_library().svo_stream_begin(self._h, gridsize, n_partitions, int(levels),
                                               None, None))

    def stream_partition(self, part: int):
        _check(load_library().svo_stream_partition(self._h, part))

    def stream_end(self):
        nn, nd, hz = C.c_uint64(0), C.c_uint64(0), C.c_int(0)
        _check(load_library().svo_stream_end(self._h, C.byref(nn), C.byref(nd), C.byref(hz)))
        return int(nn.value), int(nd.value), int(hz.value)

    def rerun(self, payload: bool, gridsize: int, memory_limit_mb: int = 2048,
              bbox=(0, 0, 0, 1, 1, 1), color_mode: int = 0, levels: bool = False) -> RunInfo:
        """The whole pipeline again on the mesh already resident in HBM."""
        b = np.asarray(bbox, np.float32)
        info = RunInfo()
        _check(load_library().svo_run(self._h, None, 0, int(payload), 0, _ptr(b), gridsize,
                                      memory_limit_mb, color_mode, int(levels), C.byref(info)))
        self.last_info = info
        return info

    def batch_run(self, jobs) -> tuple[list, RunInfo]:
        """Batched form: jobs is a list of dicts (see fill_job). Returns (per-job JobResult
        list, totals)."""
        n = len(jobs)
        arr = (Job * n)()
        keep = [fill_job(arr[i], j) for i, j in enumerate(jobs)]
        res = (JobResult * n)()
        info = RunInfo()
        _check(load_library().svo_batch_run(self._h, arr, n, res, C.byref(info)))
        self.last_info = info
        return list(res), info

    def autotune(self, jobs) -> tuple:
        """Pick the voxelizer knobs (split threshold, grid sizes) on this batch."""
        n = len(jobs)
        arr = (Job * n)()
        keep = [fill_job(arr[i], j) for i, j in enumerate(jobs)]  # noqa: F841
        out = (C.c_int * 4)()
        _check(load_library().svo_ctx_autotune(self._h, arr, n, C.byref(out)))
        return tuple(out)

    def get_tuning(self) -> tuple:
        """(small_cols, direct_cols, tile_blocks, large_blocks)"""
        out = (C.c_int * 4)()
        _check(load_library().svo_ctx_get_tuning(self._h, C.byref(out)))
        return tuple(out)

    def set_knobs(self, small_cols: int, direct_cols: int, tile_blocks: int, large_blocks: int):
        k = (C.c_int * 4)(small_cols, direct_cols, tile_blocks, large_blocks)
        _check(load_library().svo_ctx_set_knobs(self._h, C.byref(k)))

    def nodes_range(self, first: int, count: int) -> np.ndarray:
        out = np.empty(count, NODE_DTYPE)
        _check(load_library().svo_copy_nodes(self._h, _ptr(out), first, count))
        return out

    def copy_nodes_to(self, dst_ptr: int, first: int, count: int):
        """Node records [first, first+count) into device (or host) memory at dst_ptr."""
        _check(load_library().svo_copy_nodes(self._h, C.c_void_p(dst_ptr), first, count))

    def copy_data_to(self, dst_ptr: int, first: int, count: int):
        _check(load_library().svo_copy_data(self._h, C.c_void_p(dst_ptr), first, count))

    def data_range(self, first: int, count: int) -> np.ndarray:
        out = np.empty(count, DATA_DTYPE)
        _check(load_library().svo_copy_data(self._h, _ptr(out), first, count))
        return out

    def set_kernel_timing(self, on: bool):
        _check(load_library().svo_ctx_set_kernel_timing(self._h, int(on)))

    def kernel_timing(self, name: str):
        """(launches, total milliseconds) of one instrumented kernel."""
        n, ms = C.c_uint64(0), C.c_double(0)
        _check(load_library().svo_kernel_timing(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    def reset_kernel_timing(self):
        _check(load_library().svo_kernel_timing_reset(self._h))

    # -- results ----------------------------------------------------------------
    def voxel_count(self) -> int:
        return int(load_library().svo_voxel_count(self._h))

    def copy_codes_to(self, dst_ptr: int, first: int, count: int):
        """Copy sorted voxel codes [first, first + count) to a device (or host) address."""
        _check(load_library().svo_copy_codes(self._h, C.c_void_p(dst_ptr), first, count))

    def codes(self) -> np.ndarray:
        n = self.voxel_count()
        out = np.empty(n, np.uint64)
        _check(load_library().svo_copy_codes(self._h, _ptr(out), 0, n))
        return out

    def nodes(self, n_nodes: int) -> np.ndarray:
        out = np.empty(n_nodes, NODE_DTYPE)
        _check(load_library().svo_copy_nodes(self._h, _ptr(out), 0, n_nodes))
        return out

    def data(self, n_data: int) -> np.ndarray:
        out = np.empty(n_data, DATA_DTYPE)
        _check(load_library().svo_copy_data(self._h, _ptr(out), 0, n_data))
        return out
