"""Streamed (out-of-core) build: the octree assembled partition by partition
(OctreeBuilder.cpp:4-168 fed by main.cpp:329-387), against the oracle.

* svo_stream_* over voxel sets added partition by partition (svo_add_voxels): every
  partition count, empty partitions, a partition per voxel, -levels, the hazard voxel in
  the last partition — node and data bytes equal to the oracle's streaming builder fed the
  whole set.
* svo_run_stream with a Python sink, one and several partitions, binary, payload and
  -levels: the byte streams the sink receives equal the oracle's files.
* The 5x10^7-triangle sphere soup's shape at reduced size (-s 512 -l 16: 8 partitions)
  against the oracle, and at full size (-s 8192 -l 65536: 8 partitions of 4096^3) with
  size-independent properties and partition-count invariance (-l 8192: 64 partitions
  give the same bytes, since the unitlength 1/8192 is a power of two).
"""
import os
import time

import numpy as np
import pytest

import oracle
from ooc_svo_builder_amd import SvoContext, generators
from ooc_svo_builder_amd.formats import DATA_DTYPE

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _random_codes(rng, G, n, clustered=True):
    if clustered:  # runs of voxels, so that groups fill up
        starts = rng.integers(0, G ** 3, n // 8 + 1)
        codes = np.unique((starts[:, None] + np.arange(8)[None, :]).reshape(-1) % (G ** 3))
    else:
        codes = np.unique(rng.integers(0, G ** 3, n))
    return codes.astype(np.uint64)


def _payload_records(rng, codes):
    r = np.zeros(len(codes), DATA_DTYPE)
    r["morton"] = codes
    r["color"] = rng.uniform(0, 1, (len(codes), 3)).astype(np.float32)
    n = rng.normal(size=(len(codes), 3)).astype(np.float32)
    r["normal"] = n
    return r


def _stream_set(ctx, codes, recs, G, P, levels):
    """Feed a voxel set partition by partition through the streamed build (records
    collected in HBM)."""
    payload = recs is not None
    ctx.load_mesh(np.zeros((0, 21 if payload else 9), np.float32), payload)
    ctx.stream_begin(G, P, levels)
    span = G ** 3 // P
    part = codes // np.uint64(span)
    for p in np.unique(part):
        sel = part == p
        ctx.add_voxels(codes[sel], None if recs is None else recs[sel])
        ctx.stream_partition(int(p))
    return ctx.stream_end()


@pytest.mark.parametrize("G,P,payload,levels", [
    (16, 8, False, False), (16, 8, True, False), (32, 64, False, False),
    (32, 512, True, False), (8, 512, False, False), (8, 512, True, True),
    (32, 8, True, True), (64, 64, False, True), (64, 1, True, False), (64, 1, False, True)])
def test_stream_sets_match_oracle(gpu_ctx, G, P, payload, levels):
    _say("stream sets G=%d P=%d payload=%d levels=%d" % (G, P, payload, levels))
    rng = np.random.default_rng(G * 1000 + P + 7 * payload + 3 * levels)
    for trial in range(3):
        codes = _random_codes(rng, G, max(8, G ** 3 // (4 + 20 * trial)), clustered=trial != 1)
        if trial == 2:  # whole partitions left empty
            span = G ** 3 // P
            codes = codes[(codes // np.uint64(span)) % np.uint64(3) != 1]
        recs = _payload_records(rng, codes) if payload else None
        nn, nd, hz = _stream_set(gpu_ctx, codes, recs, G, P, levels)
        nodes_w, data_w, hz_w = oracle.build_octree(recs if payload else codes, G, payload,
                                                    levels=levels)
        assert (nn, nd, hz) == (len(nodes_w), len(data_w), hz_w), trial
        assert gpu_ctx.nodes(nn).tobytes() == nodes_w.tobytes(), trial
        assert gpu_ctx.data(nd).tobytes() == data_w.tobytes(), trial


@pytest.mark.parametrize("payload", [False, True])
def test_stream_hazard_and_empty(gpu_ctx, payload):
    """The last voxel at the second-to-last Morton code (in the last partition), an empty
    set, and a set in the first partition only."""
    _say("stream hazard and empty payload=%d" % payload)
    rng = np.random.default_rng(5)
    G, P = 16, 8
    for codes in (np.array([3, 100, G ** 3 - 2], np.uint64), np.zeros(0, np.uint64),
                  np.array([0, 1, 2, 9], np.uint64)):
        recs = _payload_records(rng, codes) if payload else None
        nn, nd, hz = _stream_set(gpu_ctx, codes, recs, G, P, False)
        nodes_w, data_w, hz_w = oracle.build_octree(recs if payload else codes, G, payload)
        assert (nn, nd, hz) == (len(nodes_w), len(data_w), hz_w)
        assert gpu_ctx.nodes(nn).tobytes() == nodes_w.tobytes()
        assert gpu_ctx.data(nd).tobytes() == data_w.tobytes()


def test_stream_order_is_checked(gpu_ctx):
    gpu_ctx.load_mesh(np.zeros((0, 9), np.float32), False)
    gpu_ctx.stream_begin(16, 8)
    gpu_ctx.stream_partition(3)
    with pytest.raises(Exception):
        gpu_ctx.stream_partition(2)
    assert gpu_ctx.stream_end()[0] == 1  # the empty tree: its null root


class Collect:
    def __init__(self):
        self.parts = {0: [], 1: []}
        self.calls = 0

    def __call__(self, kind, buf):
        self.calls += 1
        self.parts[kind].append(bytes(buf))

    def bytes(self, kind):
        return b"".join(self.parts[kind])


@pytest.mark.parametrize("payload,G,L,mode,levels", [
    (False, 128, 2048, 0, False), (False, 256, 2, 0, False), (True, 256, 2, 0, False),
    (True, 128, 2048, 2, False), (True, 256, 2, 3, True), (False, 512, 2, 0, True)])
def test_run_stream_sink_matches_oracle(gpu_ctx, payload, G, L, mode, levels):
    _say("run_stream payload=%d G=%d L=%d mode=%d levels=%d" % (payload, G, L, mode, levels))
    if payload:
        tris = generators.terrain(n=65, seed=31, payload=True)
    else:
        rng = np.random.default_rng(32)
        tris = np.concatenate([generators.icosphere(subdivisions=3, radius=0.3),
                               generators.sphere_soup(3000, n_spheres=5, seed=33,
                                                      median_edge=3.0 / G),
                               rng.uniform(0, 1, (50, 9)).astype(np.float32)])
    sink = Collect()
    info = gpu_ctx.run_stream(tris, payload, G, sink, L, color_mode=mode, levels=levels)
    nodes_w, data_w, hz_w, np_w, nf_w = oracle.pipeline(tris, payload, G, L, color_mode=mode,
                                                        levels=levels)
    assert (info.n_partitions, info.nfilled, info.hazard) == (np_w, nf_w, hz_w)
    assert (info.n_nodes, info.n_data) == (len(nodes_w), len(data_w))
    assert sink.bytes(0) == nodes_w.tobytes()
    assert sink.bytes(1) == data_w.tobytes()
    # the in-memory run (streamed in HBM for several partitions) gives the same bytes
    info2 = gpu_ctx.run(tris, payload, G, L, color_mode=mode, levels=levels)
    assert gpu_ctx.nodes(info2.n_nodes).tobytes() == nodes_w.tobytes()
    assert gpu_ctx.data(info2.n_data).tobytes() == data_w.tobytes()


def test_run_stream_sink_error_stops(gpu_ctx):
    _say("run_stream sink error")

    def bad(kind, buf):
        raise IOError("disk full")
    tris = generators.icosphere(subdivisions=3)
    with pytest.raises(IOError):
        gpu_ctx.run_stream(tris, False, 64, bad)
    info = gpu_ctx.run(tris, False, 64)  # the context is usable afterwards
    assert info.n_nodes > 0


def test_sphere_soup_reduced_vs_oracle(gpu_ctx):
    """The out-of-core workload's shape at -s 512 -l 16: 8 partitions of 256^3, edges
    log-normal around 4 voxels."""
    _say("sphere soup 400k at 512, 8 partitions, vs oracle")
    tris = generators.sphere_soup(400_000, median_edge=4.0 / 512, workers=4)
    info = gpu_ctx.run(tris, False, 512, 16)
    nodes_w, data_w, hz_w, np_w, nf_w = oracle.pipeline(tris, False, 512, 16)
    assert info.n_partitions == np_w == 8
    assert (info.nfilled, info.hazard) == (nf_w, hz_w)
    assert gpu_ctx.nodes(info.n_nodes).tobytes() == nodes_w.tobytes()
    assert gpu_ctx.data(info.n_data).tobytes() == data_w.tobytes()


def _node_properties(ctx, n_nodes, nfilled):
    """Size-independent checks of a geometry-only octree in HBM, in chunks: leaves are
    the voxels and point at the white record, internal nodes carry data 0, children
    precede their parent (post-order), and every node but the root is some node's child
    exactly once in total count (sum of child counts = n_nodes - 1)."""
    import torch
    chunk = 1 << 27
    leaves = children = 0
    buf = torch.empty(chunk * 24, dtype=torch.uint8, device="cuda:0")
    for s in range(0, n_nodes, chunk):
        k = min(chunk, n_nodes - s)
        ctx.copy_nodes_to(buf.data_ptr(), s, k)
        w = buf[:k * 24].view(torch.int64).view(k, 3)
        offs = buf[:k * 24].view(torch.int8).view(k, 24)[:, 16:]
        leaf = (w[:, 2] == -1)
        leaves += int(leaf.sum())
        assert bool((w[leaf, 0] == 1).all())
        assert bool((w[~leaf, 0] == 0).all())
        nch = (offs >= 0).sum(dim=1)
        children += int(nch[~leaf].sum())
        idx = torch.arange(s, s + k, device="cuda:0", dtype=torch.int64)
        last = w[:, 1] + nch.to(torch.int64) - 1
        assert bool((last[~leaf] < idx[~leaf]).all())
        assert bool((w[~leaf, 1] >= 0).all())
    assert leaves == nfilled
    assert children == n_nodes - 1


def _partition_codes(ctx, d_tris, G, P, parts):
    """The sorted voxel codes of partitions 'parts' of a P-partition run, on the device
    (voxelize only: svo_partition, then svo_voxelize per partition)."""
    import torch
    ctx.load_mesh(d_tris, False)
    u = np.float32(1.0) / np.float32(G)
    counts = ctx.partition(u, G, P)
    span = G ** 3 // P
    for p in parts:
        if counts[p]:
            ctx.voxelize(p, p * span, (p + 1) * span, u, G)
    n = ctx.voxel_count()
    out = torch.empty(n, dtype=torch.int64, device="cuda:0")
    if n:
        ctx.copy_codes_to(out.data_ptr(), 0, n)
    return out


def test_sphere_soup_full_size_8_partitions():
    """-s 8192 -l 65536 on 5x10^7 triangles: 8 partitions of 4096^3 (64 GiB dense grid
    each), streamed build in HBM, with size-independent properties of its octree.

    Against 64 partitions (-l 8192): the reference clamps a triangle's grid bbox to the
    partition (voxelizer.cpp:198-204), so a triangle whose bbox starts exactly on a
    partition face is also tested against the last voxel layer of the partition below
    (trunc gives the face index, the clamp pulls it to the layer under it) and can mark
    voxels there that a coarser split never tests. With u = 1/8192 (exact), every voxel of
    the 8-partition run is therefore a voxel of the 64-partition run, and each extra voxel
    lies in the last layer of a 2048^3 cube that is not the last layer of a 4096^3 one."""
    import torch
    workers = max(1, min(16, os.cpu_count() or 1))
    t = time.time()
    tris = generators.sphere_soup(50_000_000, workers=workers)
    _say("soup of 5e7 triangles generated in %.1f s" % (time.time() - t))
    d_tris = torch.from_numpy(tris).cuda()
    del tris
    G = 8192
    with SvoContext(0) as ctx:
        t = time.time()
        info = ctx.run(d_tris, False, G, 65536)
        _say("P=%d: %d voxels, %d nodes in %.1f s (partition %.1f, voxelize %.1f, build %.1f ms)"
             % (info.n_partitions, info.nfilled, info.n_nodes, time.time() - t,
                info.ms_partition, info.ms_voxelize, info.ms_build))
        assert info.n_partitions == 8 and info.n_data == 2 and info.nfilled > 10 ** 8
        _node_properties(ctx, info.n_nodes, info.nfilled)
        info64 = ctx.run(d_tris, False, G, 8192)
        _say("P=%d: %d voxels, %d nodes" % (info64.n_partitions, info64.nfilled, info64.n_nodes))
        assert info64.n_partitions == 64 and info64.nfilled >= info.nfilled
        extra_total = 0
        for i in range(8):
            a = _partition_codes(ctx, d_tris, G, 8, [i])
            b = _partition_codes(ctx, d_tris, G, 64, range(8 * i, 8 * i + 8))
            pos = torch.searchsorted(b, a)
            assert bool((pos < len(b)).all()) and bool((b[pos.clamp(max=len(b) - 1)] == a).all()), i
            keep = torch.ones(len(b), dtype=torch.bool, device="cuda:0")
            keep[pos] = False
            extra = b[keep].cpu().numpy().astype(np.uint64)
            extra_total += len(extra)
            for code in extra:
                xyz = [sum(((int(code) >> (3 * k + a_)) & 1) << k for k in range(21))
                       for a_ in range(3)]
                assert any(c % 2048 == 2047 and c % 4096 != 4095 for c in xyz), (i, xyz)
            del a, b, pos, keep
        assert info.nfilled + extra_total == info64.nfilled
        _say("64 partitions: %d extra voxels, all on the partition faces" % extra_total)
