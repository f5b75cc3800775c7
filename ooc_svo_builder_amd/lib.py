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

SVO_OK = 0
SVO_ERR_INVALID = 1
SVO_ERR_DEVICE = 2
SVO_ERR_NOMEM = 3
SVO_ERR_UNSUPPORTED = 4

COLOR_MODES = {"model": 0, "fixed": 1, "linear": 2, "normal": 3}

# Every function declared in include/svo_amd.h, with (restype, argtypes).
_vp, _u64, _i32, _f32 = C.c_void_p, C.c_uint64, C.c_int, C.c_float


class RunInfo(C.Structure):
    _fields_ = [("n_partitions", C.c_uint64), ("nfilled", C.c_uint64),
                ("n_nodes", C.c_uint64), ("n_data", C.c_uint64), ("hazard", C.c_int),
                ("ms_partition", C.c_float), ("ms_voxelize", C.c_float),
                ("ms_build", C.c_float)]


SIGNATURES = {
    "svo_last_error": (C.c_char_p, []),
    "svo_version": (C.c_char_p, []),
    "svo_ctx_create": (_i32, [_i32, C.POINTER(_vp)]),
    "svo_ctx_destroy": (None, [_vp]),
    "svo_ctx_set_stream": (_i32, [_vp, _vp]),
    "svo_ctx_stream": (_vp, [_vp]),
    "svo_estimate_partitions": (_u64, [_u64, _u64]),
    "svo_mesh_load": (_i32, [_vp, _vp, _u64, _i32, _i32]),
    "svo_partition": (_i32, [_vp, _f32, _u64, _u64, _vp]),
    "svo_partition_indices": (_i32, [_vp, _u64, _vp]),
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


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


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

    # -- stages -----------------------------------------------------------------
    def load_mesh(self, records, payload: bool):
        """records: (N, 9|21) float32 numpy array, or a torch tensor on the device."""
        L = load_library()
        if hasattr(records, "data_ptr"):
            n = int(records.shape[0])
            on_dev = 1 if records.is_cuda else 0
            _check(L.svo_mesh_load(self._h, C.c_void_p(records.data_ptr()), n, int(payload), on_dev))
        else:
            r = np.ascontiguousarray(records, np.float32)
            _check(L.svo_mesh_load(self._h, _ptr(r), r.shape[0], int(payload), 0))
        self.payload = bool(payload)

    def partition(self, unitlength: float, gridsize: int, n_partitions: int) -> np.ndarray:
        counts = np.zeros(n_partitions, np.uint64)
        _check(load_library().svo_partition(self._h, float(np.float32(unitlength)), gridsize,
                                            n_partitions, _ptr(counts)))
        return counts

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
        L = load_library()
        b = np.asarray(bbox, np.float32)
        info = RunInfo()
        if hasattr(records, "data_ptr"):
            n = int(records.shape[0])
            _check(L.svo_run(self._h, C.c_void_p(records.data_ptr()), n, int(payload),
                             1 if records.is_cuda else 0, _ptr(b), gridsize, memory_limit_mb,
                             color_mode, int(levels), C.byref(info)))
        else:
            r = np.ascontiguousarray(records, np.float32)
            _check(L.svo_run(self._h, _ptr(r), r.shape[0], int(payload), 0, _ptr(b), gridsize,
                             memory_limit_mb, color_mode, int(levels), C.byref(info)))
        self.payload = bool(payload)
        self.last_info = info
        return info

    def rerun(self, payload: bool, gridsize: int, memory_limit_mb: int = 2048,
              bbox=(0, 0, 0, 1, 1, 1), color_mode: int = 0, levels: bool = False) -> RunInfo:
        """The whole pipeline again on the mesh already resident in HBM."""
        b = np.asarray(bbox, np.float32)
        info = RunInfo()
        _check(load_library().svo_run(self._h, None, 0, int(payload), 0, _ptr(b), gridsize,
                                      memory_limit_mb, color_mode, int(levels), C.byref(info)))
        self.last_info = info
        return info

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
