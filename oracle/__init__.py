"""ORACLE — test infrastructure, not product code.

ctypes binding of the CPU restatement in oracle/svo_oracle.cpp (built by
oracle/Makefile into oracle/_bin/). Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module, as the checker and as the timed
CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_bin", "liboracle.so")
CLI_PAYLOAD = os.path.join(HERE, "_bin", "svo_oracle")
CLI_BINARY = os.path.join(HERE, "_bin", "svo_oracle_binary")

_lib = None


def build(force: bool = False) -> None:
    """Compile the oracle with its Makefile (g++ only)."""
    cmd = ["make", "-C", HERE, "-j4"]
    if force:
        cmd.insert(3, "-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, sz, u64, f32, i32 = C.c_void_p, C.c_size_t, C.c_uint64, C.c_float, C.c_int
        L.oracle_morton_encode_batch.argtypes = [vp, vp, vp, sz, vp]
        L.oracle_morton_decode_batch.argtypes = [vp, sz, vp, vp, vp]
        L.oracle_estimate_partitions.argtypes = [sz, sz]
        L.oracle_estimate_partitions.restype = sz
        L.oracle_partition.argtypes = [vp, sz, i32, vp, sz, sz, vp, vp]
        L.oracle_voxelize.argtypes = [vp, sz, i32, u64, u64, f32, f32]
        L.oracle_voxelize.restype = vp
        L.oracle_vox_count.argtypes = [vp]
        L.oracle_vox_count.restype = sz
        L.oracle_vox_used_data.argtypes = [vp]
        L.oracle_vox_used_data.restype = i32
        L.oracle_vox_codes.argtypes = [vp, vp]
        L.oracle_vox_records.argtypes = [vp, vp]
        L.oracle_vox_free.argtypes = [vp]
        L.oracle_build.argtypes = [vp, vp, sz, sz, i32, i32]
        L.oracle_build.restype = vp
        for f in ("oracle_build_nnodes", "oracle_build_ndata"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = sz
        L.oracle_build_hazard.argtypes = [vp]
        L.oracle_build_hazard.restype = i32
        L.oracle_build_copy.argtypes = [vp, vp, vp]
        L.oracle_build_free.argtypes = [vp]
        L.oracle_pipeline.argtypes = [vp, sz, i32, vp, sz, sz, f32, i32, i32,
                                      C.POINTER(sz), C.POINTER(sz)]
        L.oracle_pipeline.restype = vp
        L.oracle_cli_main.argtypes = [i32, C.POINTER(C.c_char_p), i32]
        L.oracle_cli_main.restype = i32
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def morton_encode(x, y, z) -> np.ndarray:
    x = np.ascontiguousarray(x, np.uint32)
    y = np.ascontiguousarray(y, np.uint32)
    z = np.ascontiguousarray(z, np.uint32)
    out = np.empty(x.shape, np.uint64)
    lib().oracle_morton_encode_batch(_p(x), _p(y), _p(z), x.size, _p(out))
    return out


def morton_decode(m):
    m = np.ascontiguousarray(m, np.uint64)
    x = np.empty(m.shape, np.uint32)
    y = np.empty(m.shape, np.uint32)
    z = np.empty(m.shape, np.uint32)
    lib().oracle_morton_decode_batch(_p(m), m.size, _p(x), _p(y), _p(z))
    return x, y, z


def estimate_partitions(gridsize: int, memory_limit_mb: int) -> int:
    return int(lib().oracle_estimate_partitions(gridsize, memory_limit_mb))


def partition(tris: np.ndarray, payload: bool, bbox, gridsize: int, n_partitions: int):
    """Per-partition triangle counts and in-order index lists."""
    t = np.ascontiguousarray(tris, np.float32)
    b = np.asarray(bbox, np.float32)
    counts = np.zeros(n_partitions, np.uint64)
    lib().oracle_partition(_p(t), t.shape[0], int(payload), _p(b), gridsize, n_partitions,
                           _p(counts), None)
    lists = np.zeros(int(counts.sum()), np.uint64)
    lib().oracle_partition(_p(t), t.shape[0], int(payload), _p(b), gridsize, n_partitions,
                           _p(counts), _p(lists))
    out, o = [], 0
    for c in counts:
        out.append(lists[o:o + int(c)])
        o += int(c)
    return counts, out


def voxelize(tris: np.ndarray, payload: bool, morton_start: int, morton_end: int,
             unitlength: float, sparseness: float = 0.10):
    """One partition through voxelize_schwarz_method. Returns (codes in emission
    order, records or None); records are 32 B VoxelData in emission order."""
    t = np.ascontiguousarray(tris, np.float32)
    L = lib()
    h = L.oracle_voxelize(_p(t), t.shape[0], int(payload), morton_start, morton_end,
                          float(np.float32(unitlength)), float(np.float32(sparseness)))
    try:
        n = L.oracle_vox_count(h)
        codes = np.empty(n, np.uint64)
        L.oracle_vox_codes(h, _p(codes))
        recs = None
        if payload:
            from ooc_svo_builder_amd.formats import DATA_DTYPE
            recs = np.empty(n, DATA_DTYPE)
            L.oracle_vox_records(h, _p(recs))
        return codes, recs
    finally:
        L.oracle_vox_free(h)


def _take_build(h):
    from ooc_svo_builder_amd.formats import DATA_DTYPE, NODE_DTYPE
    L = lib()
    try:
        nn = L.oracle_build_nnodes(h)
        nd = L.oracle_build_ndata(h)
        nodes = np.empty(nn, NODE_DTYPE)
        data = np.empty(nd, DATA_DTYPE)
        L.oracle_build_copy(h, _p(nodes), _p(data))
        return nodes, data, int(L.oracle_build_hazard(h))
    finally:
        L.oracle_build_free(h)


def build_octree(codes_or_records, gridsize: int, payload: bool, levels: bool = False):
    """OctreeBuilder fed with sorted codes (binary) or sorted records (payload).
    Returns (nodes, data, hazard)."""
    L = lib()
    if payload:
        r = np.ascontiguousarray(codes_or_records)
        h = L.oracle_build(None, _p(r), len(r), gridsize, 1, int(levels))
    else:
        c = np.ascontiguousarray(codes_or_records, np.uint64)
        h = L.oracle_build(_p(c), None, len(c), gridsize, 0, int(levels))
    return _take_build(h)


def pipeline(tris: np.ndarray, payload: bool, gridsize: int, memory_limit_mb: int = 2048,
             bbox=(0, 0, 0, 1, 1, 1), sparseness: float = 0.10, color_mode: int = 0,
             levels: bool = False):
    """main.cpp in memory. Returns (nodes, data, hazard, n_partitions, nfilled)."""
    t = np.ascontiguousarray(tris, np.float32)
    b = np.asarray(bbox, np.float32)
    npart = C.c_size_t(0)
    nfill = C.c_size_t(0)
    h = lib().oracle_pipeline(_p(t), t.shape[0], int(payload), _p(b), gridsize,
                              memory_limit_mb, float(np.float32(sparseness)), color_mode,
                              int(levels), C.byref(npart), C.byref(nfill))
    nodes, data, hz = _take_build(h)
    return nodes, data, hz, int(npart.value), int(nfill.value)
