"""Pins the oracle's streaming OctreeBuilder (OctreeBuilder.cpp:4-168, OctreeBuilder.h:50-91)
against an independent formulation of its output: the octree levels
S_l = unique(S_{l+1} >> 3), internal nodes emitted in post-order (completion order:
by the end of their Morton range, deeper first on ties), each emitting its non-null
children in child order, root last. Also checks parse-back invariants.
"""
import itertools

import numpy as np
import pytest

import oracle
from ooc_svo_builder_amd.formats import DATA_DTYPE, NODE_DTYPE


def octree_by_definition(codes, G, payload, records=None):
    D = G.bit_length() - 1
    codes = sorted(int(c) for c in codes)
    nodes = []
    if not codes:
        nodes.append((0, 0, [-1] * 8))
    else:
        levels = {D: codes}
        for l in range(D - 1, -1, -1):
            levels[l] = sorted(set(c >> 3 for c in levels[l + 1]))
        children = {}
        for l in range(1, D + 1):
            for c in levels[l]:
                children.setdefault((l - 1, c >> 3), []).append(c)
        internal = [(l, p) for l in range(D) for p in levels[l]]
        internal.sort(key=lambda lp: ((lp[1] + 1) * 8 ** (D - lp[0]), -lp[0]))
        order = []
        pos = {}
        for (l, p) in internal:
            for c in sorted(children[(l, p)]):
                pos[(l + 1, c)] = len(order)
                order.append((l + 1, c))
        pos[(0, 0)] = len(order)
        order.append((0, 0))
        rank = {c: i for i, c in enumerate(codes)}
        for (l, c) in order:
            if l == D:
                nodes.append(((1 + rank[c]) if payload else 1, 0, [-1] * 8))
            else:
                kids = sorted(children[(l, c)])
                offs = [-1] * 8
                for r, k in enumerate(kids):
                    offs[k & 7] = r
                nodes.append((0, pos[(l + 1, kids[0])], offs))
    arr = np.zeros(len(nodes), NODE_DTYPE)
    for i, (d, cb, offs) in enumerate(nodes):
        arr[i]["data"] = d
        arr[i]["children_base"] = cb
        arr[i]["offsets"] = offs
    return arr


def traverse_leaves(nodes, D):
    """Walk from the root (last record) and return the leaf Morton codes."""
    out = []
    stack = [(len(nodes) - 1, 0, 0)]
    while stack:
        idx, level, prefix = stack.pop()
        n = nodes[idx]
        offs = n["offsets"]
        if level == D:
            if n["data"] != 0:
                out.append(prefix)
            continue
        for k in range(8):
            if offs[k] >= 0:
                stack.append((int(n["children_base"]) + int(offs[k]), level + 1, prefix * 8 + k))
    return sorted(out)


def test_exhaustive_2cube():
    """Every subset of the 8 voxels of a 2^3 grid."""
    for bits in range(256):
        codes = [k for k in range(8) if bits >> k & 1]
        nodes, data, hz = oracle.build_octree(np.array(codes, np.uint64), 2, False)
        want = octree_by_definition(codes, 2, False)
        assert nodes.tobytes() == want.tobytes(), codes
        assert hz == (1 if codes and codes[-1] == 6 else 0)
        assert len(data) == 2


@pytest.mark.parametrize("G,n", [(4, 5), (4, 40), (8, 100), (16, 700), (64, 3000), (64, 30)])
def test_random_sets(G, n):
    rng = np.random.default_rng(G * 1000 + n)
    for trial in range(6):
        codes = np.unique(rng.choice(G ** 3, size=n, replace=False)).astype(np.uint64)
        nodes, data, hz = oracle.build_octree(codes, G, False)
        want = octree_by_definition(codes, G, False)
        assert nodes.tobytes() == want.tobytes()
        D = G.bit_length() - 1
        assert traverse_leaves(nodes, D) == [int(c) for c in codes]
        levels = [set(int(c) for c in codes)]
        for _ in range(D):
            levels.append(set(c >> 3 for c in levels[-1]))
        assert len(nodes) == sum(len(s) for s in levels[:-1]) + 1


def test_clustered_sets():
    """Dense clusters and runs (full octants collapse only when all 8 are present)."""
    G = 32
    rng = np.random.default_rng(5)
    codes = set(range(0, 512))  # a full 8^3 block
    codes |= set(int(c) for c in rng.integers(4096, 9000, 800))
    codes |= set(range(G ** 3 - 64, G ** 3))  # the last 4^3 block, incl. the max code
    codes = np.array(sorted(codes), np.uint64)
    nodes, data, hz = oracle.build_octree(codes, G, False)
    assert nodes.tobytes() == octree_by_definition(codes, G, False).tobytes()
    assert hz == 0


def test_payload_data_order():
    G = 16
    rng = np.random.default_rng(2)
    codes = np.unique(rng.choice(G ** 3, 500, replace=False)).astype(np.uint64)
    recs = np.zeros(len(codes), DATA_DTYPE)
    recs["morton"] = codes
    recs["color"] = rng.uniform(0, 1, (len(codes), 3))
    recs["normal"] = rng.normal(0, 1, (len(codes), 3))
    nodes, data, hz = oracle.build_octree(recs, G, True)
    assert nodes.tobytes() == octree_by_definition(codes, G, True).tobytes()
    assert len(data) == len(codes) + 1
    assert data[0].tobytes() == bytes(32)
    assert data[1:].tobytes() == recs.tobytes()


def test_empty_and_binary_data_records():
    nodes, data, hz = oracle.build_octree(np.zeros(0, np.uint64), 8, False)
    assert len(nodes) == 1 and nodes[0]["data"] == 0 and nodes[0]["children_base"] == 0
    assert list(nodes[0]["offsets"]) == [-1] * 8
    # OctreeBuilder.cpp:25-29: a null record then one white record
    assert data[0].tobytes() == bytes(32)
    assert data[1]["morton"] == 0
    assert list(data[1]["color"]) == [1.0, 1.0, 1.0]
    assert list(data[1]["normal"]) == [0.0, 0.0, 0.0]
    nodes, data, hz = oracle.build_octree(np.zeros(0, DATA_DTYPE), 8, True)
    assert len(nodes) == 1 and len(data) == 1


def test_single_voxel_grid():
    nodes, data, hz = oracle.build_octree(np.array([0], np.uint64), 1, False)
    assert len(nodes) == 1 and nodes[0]["data"] == 1 and hz == 0
    nodes, data, hz = oracle.build_octree(np.zeros(0, np.uint64), 1, False)
    assert len(nodes) == 1 and hz == 1  # the reference reads an empty buffer here


def _post_order(nodes, D):
    """(record index, level) of every node below the root, children in child order, each
    node after its subtree (the order in which the streaming builder completes nodes)."""
    out = []

    def visit(idx, level):
        n = nodes[idx]
        if level < D:
            for k in range(8):
                if n["offsets"][k] >= 0:
                    visit(int(n["children_base"]) + int(n["offsets"][k]), level + 1)
        out.append((idx, level))
    visit(len(nodes) - 1, 0)
    return out


def _f32_average(kids):
    """OctreeBuilder.cpp:82-99 in float32: the 8 child caches summed in child order (None =
    null child, adds zeros), divided by the non-null count, the normal normalized."""
    f = np.float32
    c = [f(0), f(0), f(0)]
    nrm = [f(0), f(0), f(0)]
    notnull = f(0)
    for kid in kids:
        col, nor = (np.zeros(3, np.float32), np.zeros(3, np.float32)) if kid is None else kid
        if kid is not None:
            notnull = f(notnull + f(1))
        for a in range(3):
            c[a] = f(c[a] + col[a])
            nrm[a] = f(nrm[a] + nor[a])
    col = np.array([f(x / notnull) for x in c], np.float32)
    t = [f(x / notnull) for x in nrm]
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = f(f(1) / np.sqrt(f(f(t[0] * t[0] + t[1] * t[1]) + f(t[2] * t[2]))))
        nor = np.array([f(x * inv) for x in t], np.float32)
    return col, nor


@pytest.mark.parametrize("payload", [True, False])
def test_levels_post_order(payload):
    """-levels: data indices are post-order ranks (every node for payload, internal nodes
    only for binary after the white record), leaves keep their records, internal records
    are the child averages with morton 0; binary leaves have zero caches, so their
    ancestors' normals are NaN (written as the x86 default NaN)."""
    G, D = 16, 4
    rng = np.random.default_rng(31 + payload)
    codes = np.unique(rng.choice(G ** 3, 300, replace=False)).astype(np.uint64)
    if payload:
        recs = np.zeros(len(codes), DATA_DTYPE)
        recs["morton"] = codes
        recs["color"] = rng.uniform(0, 1, (len(codes), 3))
        recs["normal"] = rng.normal(0, 1, (len(codes), 3))
        nodes, data, hz = oracle.build_octree(recs, G, True, levels=True)
    else:
        nodes, data, hz = oracle.build_octree(codes, G, False, levels=True)
    # the node structure is the one without -levels; only internal data fields change
    plain = octree_by_definition(codes, G, payload)
    assert np.array_equal(nodes["children_base"], plain["children_base"])
    assert np.array_equal(nodes["offsets"], plain["offsets"])
    order = _post_order(nodes, D)
    internal = [(i, l) for i, l in order if l < D]
    assert len(data) == (1 + len(codes) if payload else 2) + len(internal)
    cache = {}
    rank = 0
    leaf_rank = 0
    for idx, level in order:
        n = nodes[idx]
        if level == D:
            if payload:
                assert int(n["data"]) == 1 + rank
                assert data[int(n["data"])].tobytes() == recs[leaf_rank].tobytes()
                cache[idx] = (recs[leaf_rank]["color"], recs[leaf_rank]["normal"])
                rank += 1
            else:
                assert int(n["data"]) == 1
                cache[idx] = (np.zeros(3, np.float32), np.zeros(3, np.float32))
            leaf_rank += 1
            continue
        want_idx = (1 + rank) if payload else (2 + rank)
        assert int(n["data"]) == want_idx
        rank += 1
        kids = [cache[int(n["children_base"]) + int(n["offsets"][k])] if n["offsets"][k] >= 0
                else None for k in range(8)]
        col, nor = _f32_average(kids)
        cache[idx] = (col, nor)
        rec = data[want_idx]
        assert int(rec["morton"]) == 0
        assert rec["color"].tobytes() == col.tobytes()
        if payload:
            assert rec["normal"].tobytes() == nor.tobytes()
        else:
            assert np.isnan(nor).all()
            assert rec["normal"].view(np.uint32).tolist() == [0xFFC00000] * 3
